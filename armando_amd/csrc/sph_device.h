// sph_device.h -- device-side state, parameters and pair physics of the
// WCSPH step, for the FP64 plain form (R = double) and the FP32 chosen form
// (R = float).  Included by armando.hip only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arm {

template <typename R> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };

// Scalar parameters, passed to every kernel by value.
struct KParams {
  int n, np;
  int all_rate;  // 1: walls integrate density/energy too (dynamic boundary)
  int flags, pa_sph, density_method;
  int nx, ny, ncell;
  double h, invh, mass, rho0, c0, pmin, B;  // B = rho0 c0^2 / 7
  double av_alpha, av_beta, av_e2;          // av_e2 = (h eta)^2
  double ah_g1, ah_g2, ah_e2;
  double xsph_eps, gx, gy, dt;
  double ox, oy, cell;
  double cut2;   // (2h)*(2h)
  double wnorm;  // Wendland 2D normalisation 7/(4 pi h^2)
  double C1;     // grad W = C1 (q-2)^3 (x_i - x_j), C1 = 5 wnorm / (8 h^2)
  double psi0;   // g1 h c0 (constant part of the heat coefficient)
  double g2h2;   // g2 h h
  double eps_m;  // XSPH eps * mass
  int same_e2;   // artificial viscosity and heat share (h eta)^2
  // FP32 pair body constants: m F = f_cm s^3 with s = 1 - q/2 (f_cm = -8 C1 m);
  // (beta phi - alpha c0)/2 = f_avb phi + f_ava; XSPH wnorm eps m / 2; 2 cond
  double f_cm, f_ava, f_avb, f_xs, f_cond2;
  float cellf, cut2f;
  // FP32 neighbour pre-test: a bound on (2h)^2 for r^2 computed from
  // x = cx cell + off rounded once per particle (the exact test follows)
  float pre2f;
  // FP32 roundings of the constants the FP32 pair bodies use, so that they
  // stay in scalar registers (a conversion in the kernel lands in a vector
  // register); each is (float) of the double field it is named after, apart
  // from four products: avah = f_ava h, avbh2 = f_avb h^2, mC1 = mass C1 and
  // mw = mass wnorm
  struct {
    float invh, cm, avah, avbh2, xs, ave2, ahe2, cond2, rho0, gx, gy, C1, mass, wnorm, mC1, mw;
  } kf;
  int eos;         // 0 Tait, 1 linear
  double lin_c2;   // linear EOS: p = lin_c2 rho + lin_p0
  double lin_p0;
  double lin_p0r;  // lin_p0 + lin_c2 rho0 (FP32 form: p = lin_c2 rho' + lin_p0r)
  double eta;      // dynamic viscosity (flag 32)
  double hs_q0, hs_x, hs_y, hs_k;  // heat source (flag 64): q0 exp(-k |x - c|^2) / rho
  // three dimensions: key = ix + iy nx + iz nx ny over nz layers from oz;
  // wnorm and C1 then hold the 3D Wendland constants
  int dim, nz;
  double oz, gz;
  int periodic;  // flag 256: periodic in x with period nx * cell
  double cond;   // flag 128: 4 k k/(k + k) / cp, the conduction factor on u_i - u_j
  // smoothing-length evolution (hsml.f): 0 fixed h; 1 h = 2 sqrt(m/rho);
  // 2 dh/dt = (h/2) div v.  With 1 or 2 the state carries h per particle
  // and every pair uses mh = (h_i + h_j)/2 (with_h below).
  int sle;
  double av_eta, ah_eta;
  // mirror ghosts (flag 512, ghost.f): ids >= nb (= np + nv) are the ghosts
  // of the current step; without ghosts nb = n
  int nb;
  // binned energy deposition (fluka_heat.f, Cartesian binning): map of
  // dep_n[0] dep_n[1] dep_n[2] values (W/m^3, x fastest) from dep_o in bins
  // of dep_d; applied in the steps that carry KF_DEPOSITION in flags
  const double* dep;
  int dep_n[3];
  double dep_o[3], dep_d[3];
  int dep_type;  // 0: Cartesian axes (x, y, z); 1: R-phi-z axes (r, phi in degrees, z)
  // stop at the first domain error of a fused run of steps: its is the step
  // the launch belongs to; *halt (null outside fused runs) is 0, or the
  // first step that must not run (set by the launch that found a particle
  // outside the grid, or an h that outgrew the cell)
  int its;
  int* halt;
  double contact_h;  // flag 1024: the contact distance h / hdr
};

// True in every launch of a step at or after the recorded failure.  The
// word is written only by earlier launches, so all lanes agree.
__device__ __forceinline__ bool halted(const KParams& K) {
  if (!K.halt) return false;
  const int h = *K.halt;
  return h != 0 && K.its >= h;
}

// Records that steps after this launch's step must not run (first failure wins).
__device__ __forceinline__ void raise_halt(const KParams& K) {
  if (K.halt) atomicCAS(K.halt, 0, K.its + 1);
}

// The dispatcher deals the workgroups of a launch to the eight XCDs in turn
// (workgroup b runs on XCD b mod 8), and each XCD has its own L2.  xcd_block
// renumbers them so that XCD x covers one contiguous share of the grid:
// neighbouring particles, which read the same neighbour rows, then share an
// L2.  A bijection of 0 .. gridDim.x-1 for any grid size.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_block() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb / NXCD, r = nb - q * NXCD, x = b - (b / NXCD) * NXCD;
  return x * q + (x < r ? x : r) + b / NXCD;
}

// The same for the first `count` logical blocks of a grid of at least count
// workgroups: XCD x covers one contiguous share of 0 .. count-1, the shares
// as even as the count allows; -1 for the workgroups past them.
__device__ __forceinline__ int xcd_share(int count) {
  const int b = (int)blockIdx.x;
  const int q = count / NXCD, r = count - q * NXCD, x = b - (b / NXCD) * NXCD, j = b / NXCD;
  return j < q + (x < r ? 1 : 0) ? x * q + (x < r ? x : r) + j : -1;
}

// Set in KParams::flags by the host for the steps whose time its dt lies
// within the deposition time (single_step.f:219); not a caller flag.
constexpr int KF_DEPOSITION = 1 << 24;

__host__ __device__ inline void set_kf(KParams& K) {
  K.kf.invh = (float)K.invh;
  K.kf.cm = (float)K.f_cm;
  K.kf.avah = (float)(K.f_ava * K.h);
  K.kf.avbh2 = (float)(K.f_avb * K.h * K.h);
  K.kf.xs = (float)K.f_xs;
  K.kf.ave2 = (float)K.av_e2;
  K.kf.ahe2 = (float)K.ah_e2;
  K.kf.cond2 = (float)K.f_cond2;
  K.kf.rho0 = (float)K.rho0;
  K.kf.gx = (float)K.gx;
  K.kf.gy = (float)K.gy;
  K.kf.C1 = (float)K.C1;
  K.kf.mass = (float)K.mass;
  K.kf.wnorm = (float)K.wnorm;
  K.kf.mC1 = (float)(K.mass * K.C1);
  K.kf.mw = (float)(K.mass * K.wnorm);
}

// K with the smoothing length mh: h, 1/h, the Wendland 2D constants, the
// cut-off (2 mh)^2, the artificial viscosity and heat regularisations
// (mh eta)^2 and the art_heat coefficient constants of mh.
__device__ __forceinline__ KParams with_h(const KParams& K, double mh) {
  KParams P = K;
  P.h = mh;
  P.invh = 1.0 / mh;
  P.wnorm = 7.0 / (4.0 * M_PI * mh * mh);
  P.C1 = 0.625 * P.wnorm / (mh * mh);
  P.cut2 = (2.0 * mh) * (2.0 * mh);
  P.cut2f = (float)P.cut2;
  P.av_e2 = (mh * K.av_eta) * (mh * K.av_eta);
  P.ah_e2 = (mh * K.ah_eta) * (mh * K.ah_eta);
  P.psi0 = K.ah_g1 * mh * K.c0;
  P.g2h2 = K.ah_g2 * mh * mh;
  P.f_cm = -8.0 * P.C1 * K.mass;
  P.f_xs = 0.5 * P.wnorm * K.eps_m;
  set_kf(P);
  return P;
}

// One axis of the deposition interpolation (fluka_heat.f:256-286): the bins
// on either side of the position, clamped to the map, and the local
// coordinate t in [-1, 1) between their centres.
__device__ __forceinline__ void dep_axis(double s, int n, int& lo, int& hi, double& t) {
  const int ir = (int)s;  // truncation, as Fortran int()
  double xr = s - ir;
  if (xr >= 0.5) {
    lo = ir;
    hi = ir + 1;
    xr = 2.0 * xr - 2.0;
  } else {
    lo = ir - 1;
    hi = ir;
    xr = 2.0 * xr;
  }
  lo = lo < 0 ? 0 : lo;
  hi = hi >= n ? n - 1 : hi;
  t = xr;
}

// Trilinear value of the deposition map at (x, y, z) between bin centres
// (fluka_heat.f:235-337, dim = 3, tfluka = 0), 0 outside the binning.  The
// 2D branch (:485-573) is the same at z = 0: (0 - oz)/dz is its -oz/dz, and
// z = 0 was checked to lie within the map when it was set.
//
// R-phi-z binning (dep_type 1; fluka_heat.f:341-481 in 3D, :593-704 in 2D):
// the map's axes are r = sqrt(x^2 + y^2), phi = 180/pi atan2(y, x) (pi the
// REAL(4) literal of options.inc:21) and z; the bins and weights are those
// of the Cartesian branch on these coordinates (the reference's branch
// scales z by the phi bin width and, in 3D, picks the bin pairs again after
// clamping them; neither is followed).
__device__ __forceinline__ double deposition_at(const KParams& K, double x, double y, double z) {
  const int* n = K.dep_n;
  const double* o = K.dep_o;
  const double* d = K.dep_d;
  if (K.dep_type == 1) {
    const double r = sqrt(x * x + y * y);
    const double phi = 180.0 / 3.1415927410125732 * atan2(y, x);
    x = r;
    y = phi;
  }
  if (!(x >= o[0] && y >= o[1] && z >= o[2] && x <= o[0] + n[0] * d[0] &&
        y <= o[1] + n[1] * d[1] && z <= o[2] + n[2] * d[2]))
    return 0.0;
  int i0, i1, j0, j1, k0, k1;
  double a, b, c;
  dep_axis((x - o[0]) / d[0], n[0], i0, i1, a);
  dep_axis((y - o[1]) / d[1], n[1], j0, j1, b);
  dep_axis((z - o[2]) / d[2], n[2], k0, k1, c);
  const size_t sy = (size_t)n[0], sz = (size_t)n[0] * (size_t)n[1];
  const double* m = K.dep;
  auto at = [&](int i, int j, int k) { return m[(size_t)i + (size_t)j * sy + (size_t)k * sz]; };
  double v = 0.125 * (1.0 - a) * (1.0 - b) * (1.0 - c) * at(i0, j0, k0);
  v = v + 0.125 * (1.0 + a) * (1.0 - b) * (1.0 - c) * at(i1, j0, k0);
  v = v + 0.125 * (1.0 + a) * (1.0 + b) * (1.0 - c) * at(i1, j1, k0);
  v = v + 0.125 * (1.0 - a) * (1.0 + b) * (1.0 - c) * at(i0, j1, k0);
  v = v + 0.125 * (1.0 - a) * (1.0 - b) * (1.0 + c) * at(i0, j0, k1);
  v = v + 0.125 * (1.0 + a) * (1.0 - b) * (1.0 + c) * at(i1, j0, k1);
  v = v + 0.125 * (1.0 + a) * (1.0 + b) * (1.0 + c) * at(i1, j1, k1);
  v = v + 0.125 * (1.0 - a) * (1.0 + b) * (1.0 + c) * at(i0, j1, k1);
  return v;
}

// Gaussian heat source factor exp(-k |x - c|^2) at a position.
__device__ __forceinline__ double heat_shape(const KParams& K, double x, double y) {
  const double dx = x - K.hs_x, dy = y - K.hs_y;
  return exp(-K.hs_k * (dx * dx + dy * dy));
}

// Particle state in sorted (cell, id) order.  FP64 form: xy holds the
// positions, rho the density.  FP32 form: xy holds the FP64 position
// accumulators, (cxy, off) the integer cell and FP32 in-cell offset used by
// the pair arithmetic, rho the density perturbation rho - rho0, uf a float
// copy of the FP64 energy accumulator u.
template <typename R> struct St {
  using R2 = typename V2<R>::type;
  int32_t* id;
  double2* xy;
  uint32_t* cxy;
  float2* off;
  R2* v;
  R* rho;
  double* u;
  float* uf;
  R* p;
  R* psi;    // art_heat coefficient psi_i = g1 h c + g2 h^2 (|divv| - divv)
  R* drho;   // density rate
  R* du;     // energy rate
  R2* a;     // acceleration
  R2* av;    // XSPH correction
  R* divv;   // velocity divergence
  R* rho_min;
  double* u_min;
  R2* v_min;
  uint8_t* flag;  // bit 0: predictor clamped u at 0; bit 1: EOS clamp taken
  // physical viscosity, per particle: 4 values (Txx, Tyy, Txy, tds), T the
  // shear tensor and tds = eta/(2 rho) T:T; the force uses W = eta T
  // (pa_sph 0) or eta T / rho^2 (pa_sph 1)
  R* tau;
  // per-particle smoothing length (KParams.sle != 0, FP64 in both forms) and
  // the updated one written by the force pass
  double* hs;
  double* hn;
};

__device__ __forceinline__ bool is_rate(const KParams& K, int id) {
  return K.all_rate || id < K.np;
}

// ---- positions seen by the pair arithmetic ----------------------------
struct PosD { double x, y; };
struct PosF { int cx, cy; float ox, oy; };

__device__ __forceinline__ PosD load_pos(const St<double>& S, int t) {
  const double2 q = S.xy[t];
  return {q.x, q.y};
}
__device__ __forceinline__ PosF load_pos(const St<float>& S, int t) {
  const uint32_t c = S.cxy[t];
  const float2 o = S.off[t];
  return {int(c & 0xffffu), int(c >> 16), o.x, o.y};
}

// x_i - x_j, with the neighbour sh cells (sh = 0 or -+nx across the seam of
// a periodic grid) to the right.  The module is compiled without
// floating-point contraction, so these are the exact operation sequences
// the oracle restates: (x_i - x_j) + sh cell in FP64, and
// (off_i - off_j) + (float)(c_i - c_j + sh) cell in the FP32 form.
__device__ __forceinline__ void delta(const PosD& a, const PosD& b, const KParams& K, double& dx,
                                      double& dy, int sh = 0) {
  dx = a.x - b.x;
  if (sh) dx = dx + (double)sh * K.cell;
  dy = a.y - b.y;
}
__device__ __forceinline__ void delta(const PosF& a, const PosF& b, const KParams& K, float& dx,
                                      float& dy, int sh = 0) {
  const float tx = a.ox - b.ox;
  const float ux = (float)(a.cx - b.cx + sh) * K.cellf;
  dx = tx + ux;
  const float ty = a.oy - b.oy;
  const float uy = (float)(a.cy - b.cy) * K.cellf;
  dy = ty + uy;
}

template <typename R> __device__ __forceinline__ R cutoff2(const KParams& K);
template <> __device__ __forceinline__ double cutoff2<double>(const KParams& K) { return K.cut2; }
template <> __device__ __forceinline__ float cutoff2<float>(const KParams& K) { return K.cut2f; }

__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// a*b + c in the pair bodies: one fused rounding in the FP32 form; in the
// FP64 form a rounded product and a rounded sum, the oracle's arithmetic
// (the callers order the operands so that this is the same sum as the
// reference's expression).  The neighbour criterion (dx, dy, r^2) never
// uses it, so the FP32 pair sets stay those of the unfused restatement.
__device__ __forceinline__ float fmr(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmr(double a, double b, double c) { return a * b + c; }

// Density in the pair arithmetic: absolute (FP64 form) or rho0 + perturbation.
__device__ __forceinline__ double dens(const KParams&, double r) { return r; }
__device__ __forceinline__ float dens(const KParams& K, float r) { return K.kf.rho0 + r; }

// Compensated (Kahan) accumulator for the FP32 form; plain for FP64.
template <typename R> struct Acc {
  R s = 0;
  __device__ __forceinline__ void add(R x) { s = s + x; }
  __device__ __forceinline__ void add_if(R x, bool ok) { s = ok ? s + x : s; }
  __device__ __forceinline__ R get() const { return s; }
};
template <> struct Acc<float> {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  // add(x) when ok, else nothing (both words kept), without a branch
  __device__ __forceinline__ void add_if(float x, bool ok) {
    const float y = x - c;
    const float t = s + y;
    const float cn = (t - s) - y;
    s = ok ? t : s;
    c = ok ? cn : c;
  }
  __device__ __forceinline__ float get() const { return s; }
};

// Linear EOS p = c0^2 rho + p0 from the stored density.
__device__ __forceinline__ double linear_p(const KParams& K, double rho) {
  return K.lin_c2 * rho + K.lin_p0;
}
__device__ __forceinline__ float linear_p(const KParams& K, float rp) {
  return (float)K.lin_c2 * rp + (float)K.lin_p0r;
}

// Tait pressure from the stored density (eos.f90:274).
__device__ __forceinline__ double tait(const KParams& K, double rho) {
  const double t = rho / K.rho0;
  const double t2 = t * t;
  const double t4 = t2 * t2;
  return K.rho0 * K.c0 * K.c0 / 7.0 * (t4 * t2 * t - 1.0);
}
// FP32 form: (1+x)^7 - 1 expanded in x = rho'/rho0 so that small density
// perturbations keep full relative precision.
__device__ __forceinline__ float tait(const KParams& K, float rp) {
  const float x = rp / (float)K.rho0;
  const float poly = 7.f + x * (21.f + x * (35.f + x * (35.f + x * (21.f + x * (7.f + x)))));
  return (float)K.B * (x * poly);
}

}  // namespace arm
