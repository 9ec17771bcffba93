// sph3d.h -- the WCSPH step in three dimensions for gfx950.
//
// The same routines, arguments and pair arithmetic as the 2D kernels of
// sph_kernels.h, with
//   * the hash key ix + iy nx + iz nx ny (CERNsource/link_list.f:49-99 with
//     dim = 3 is the reference's 3D grid; Code/armando.cu:3190-3317 its 27-cell
//     GPU list);
//   * the 27 neighbour cells walked as nine rows (z outer, y inner) of three
//     consecutive x-cells: in (cell, id) order each row is one contiguous
//     range [cell_start[c-1], cell_start[c+2]) of the sorted arrays;
//   * Wendland C2 in 3D: KParams.wnorm = 21/(16 pi h^3), C1 = 5 wnorm/(8 h^2)
//     (grad W = C1 (q-2)^3 (x_i - x_j) as in 2D);
//   * FP32 form: an integer cell and an FP32 in-cell offset per axis; the
//     difference of two positions per axis is
//     (off_i - off_j) + (float)(c_i - c_j) * (float)cell.  A particle keeps
//     (off_x, off_y, off_z, (float)cx) in one float4 so that a neighbour's
//     position is one 16-byte load; the y and z cell differences are those
//     of the neighbour row being walked.
// Every lane writes only its own particle: no atomics, bitwise repeatable.
#pragma once

#include "sph_kernels.h"

namespace arm {

template <typename R> struct V4;
template <> struct V4<double> { using type = double4; };
template <> struct V4<float> { using type = float4; };

// Particle state in sorted (cell, id) order (fourth vector lanes unused).
// FP64 form: xyz holds the positions, rho the density.  FP32 form: xyz holds
// the FP64 position accumulators, pf the cell + offset that the pair
// arithmetic reads, rho the perturbation rho - rho0, uf a float copy of u.
template <typename R> struct St3 {
  using R4 = typename V4<R>::type;
  int32_t* id;
  double4* xyz;
  float4* pf;
  R4* v;
  R* rho;
  double* u;
  float* uf;
  R* p;
  R* psi;
  R* drho;
  R* du;
  R4* a;
  R4* av;
  R* divv;
  R* rho_min;
  double* u_min;
  R4* v_min;
  uint8_t* flag;
};

struct Pos3D {
  double x, y, z;
};
struct Pos3F {
  float cxf, ox, oy, oz;  // cell column as float, in-cell offsets
};

__device__ __forceinline__ Pos3D load_pos3(const St3<double>& S, int t) {
  const double4 q = S.xyz[t];
  return {q.x, q.y, q.z};
}
__device__ __forceinline__ Pos3F load_pos3(const St3<float>& S, int t) {
  const float4 q = S.pf[t];
  return {q.w, q.x, q.y, q.z};
}

// The cell-difference term (float)(c_i - c_j) * cell of one axis (FP32
// form); in the row walk the y and z differences are the constants of the
// row, 1 - k % 3 and 1 - k / 3 for row k.  Unused in the FP64 form.
template <typename R> __device__ __forceinline__ R row_u3(const KParams& K, int d) {
  return (R)((float)d * K.cellf);
}
template <> __device__ __forceinline__ double row_u3<double>(const KParams&, int) { return 0.0; }

// x_i - x_j and r^2 = (dx^2 + dy^2) + dz^2, rounded step by step (no
// contraction in this module), as the oracle restates them; uy, uz are the
// row's cell terms.  The x cell columns differ by at most one, exactly
// representable, so (cxf_i - cxf_j) is the integer difference.
__device__ __forceinline__ double delta3(const Pos3D& a, const Pos3D& b, const KParams&, double,
                                         double, double& dx, double& dy, double& dz) {
  dx = a.x - b.x;
  dy = a.y - b.y;
  dz = a.z - b.z;
  double r2 = dx * dx;
  r2 = r2 + dy * dy;
  return r2 + dz * dz;
}
__device__ __forceinline__ float delta3(const Pos3F& a, const Pos3F& b, const KParams& K,
                                        float uy, float uz, float& dx, float& dy, float& dz) {
  dx = (a.ox - b.ox) + (a.cxf - b.cxf) * K.cellf;
  dy = (a.oy - b.oy) + uy;
  dz = (a.oz - b.oz) + uz;
  float r2 = dx * dx;
  r2 = r2 + dy * dy;
  return r2 + dz * dz;
}

// ------------------------------------------------------------------------
// cell keys in FP64
// ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t cell_of3(const KParams& K, double x, double y, double z,
                                             int& ix, int& iy, int& iz, bool& ok) {
  const double tx = (x - K.ox) / K.cell;
  const double ty = (y - K.oy) / K.cell;
  const double tz = (z - K.oz) / K.cell;
  ok = (tx >= 0.0) && (ty >= 0.0) && (tz >= 0.0) && (tx < (double)K.nx) &&
       (ty < (double)K.ny) && (tz < (double)K.nz);
  ix = ok ? (int)tx : 0;
  iy = ok ? (int)ty : 0;
  iz = ok ? (int)tz : 0;
  ix = ix < K.nx ? ix : K.nx - 1;
  iy = iy < K.ny ? iy : K.ny - 1;
  iz = iz < K.nz ? iz : K.nz - 1;
  return (uint32_t)ix + (uint32_t)K.nx * ((uint32_t)iy + (uint32_t)K.ny * (uint32_t)iz);
}

// FP32 form: cell + offset of an FP64 position.
__device__ __forceinline__ float4 split3(const KParams& K, double4 q) {
  int ix, iy, iz;
  bool ok;
  cell_of3(K, q.x, q.y, q.z, ix, iy, iz, ok);
  const double bx = K.ox + (double)ix * K.cell;
  const double by = K.oy + (double)iy * K.cell;
  const double bz = K.oz + (double)iz * K.cell;
  return make_float4((float)(q.x - bx), (float)(q.y - by), (float)(q.z - bz), (float)ix);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_keys(St3<R> S, KParams K, KeyOut ko,
                                                 int* __restrict__ err) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = s < K.n;
  uint32_t key = 0;
  int id = 0;
  if (valid) {
    const double4 q = S.xyz[s];
    int ix, iy, iz;
    bool ok;
    key = cell_of3(K, q.x, q.y, q.z, ix, iy, iz, ok);
    if (!ok) {
      atomicAdd(err, 1);
      raise_halt(K);
    }
    id = S.id[s];
    if constexpr (sizeof(R) == 4) S.pf[s] = split3(K, q);
  }
  emit_key(ko, valid, s, id, key);
}

// Permute every field from A to B (per-routine link_list).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_reorder_all(St3<R> A, St3<R> Bs, KParams K,
                                                        const uint32_t* __restrict__ skey,
                                                        const uint32_t* __restrict__ sval,
                                                        int* __restrict__ cstart, int fill) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = t < K.n;
  if (fill) fill_cell_start(skey, t, valid, K.n, K.ncell, cstart);
  if (!valid) return;
  const int s = (int)sval[t];
  Bs.id[t] = A.id[s];
  Bs.xyz[t] = A.xyz[s];
  if constexpr (sizeof(R) == 4) {
    Bs.pf[t] = A.pf[s];
    Bs.uf[t] = A.uf[s];
  }
  Bs.v[t] = A.v[s];
  Bs.rho[t] = A.rho[s];
  Bs.u[t] = A.u[s];
  Bs.p[t] = A.p[s];
  Bs.psi[t] = A.psi[s];
  Bs.drho[t] = A.drho[s];
  Bs.du[t] = A.du[s];
  Bs.a[t] = A.a[s];
  Bs.av[t] = A.av[s];
  Bs.divv[t] = A.divv[s];
  Bs.rho_min[t] = A.rho_min[s];
  Bs.u_min[t] = A.u_min[s];
  Bs.v_min[t] = A.v_min[s];
  Bs.flag[t] = A.flag[s];
}

// Fused reorder + leapfrog predictor (time_integration.f:136-154).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_reorder_pred(St3<R> A, St3<R> Bs, KParams K,
                                                         const uint32_t* __restrict__ skey,
                                                         const uint32_t* __restrict__ sval,
                                                         int* __restrict__ cstart, int predict,
                                                         int fill) {
  if (halted(K)) return;
  using R4 = typename V4<R>::type;
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = t < K.n;
  if (fill) fill_cell_start(skey, t, valid, K.n, K.ncell, cstart);
  if (!valid) return;
  const int s = (int)sval[t];
  const int id = A.id[s];
  Bs.id[t] = id;
  const double4 q = A.xyz[s];
  Bs.xyz[t] = q;
  if constexpr (sizeof(R) == 4) Bs.pf[t] = split3(K, q);
  R4 v = A.v[s];
  R rho = A.rho[s];
  double u = A.u[s];
  uint8_t fl = 0;
  if (predict) {
    const double hdt = K.dt / 2.0;
    if (is_rate(K, id)) {
      Bs.u_min[t] = u;
      u = u + hdt * (double)A.du[s];
      if (u < 0.0) {
        u = 0.0;
        fl = 1u;
      }
      if (K.density_method == 0) {
        Bs.rho_min[t] = rho;
        rho = rho + (R)hdt * A.drho[s];
      }
    }
    if (id < K.np) {
      const R4 a = A.a[s];
      Bs.v_min[t] = v;
      v.x = v.x + (R)hdt * a.x;
      v.y = v.y + (R)hdt * a.y;
      v.z = v.z + (R)hdt * a.z;
    }
  }
  Bs.v[t] = v;
  Bs.rho[t] = rho;
  Bs.u[t] = u;
  if constexpr (sizeof(R) == 4) Bs.uf[t] = (float)u;
  Bs.flag[t] = fl;
  if (!is_rate(K, id)) Bs.p[t] = A.p[s];
}

// ------------------------------------------------------------------------
// neighbour rows
// ------------------------------------------------------------------------
struct Rows9 {
  int t0[9], t1[9];
};

__device__ __forceinline__ Rows9 rows3_of(const KParams& K, uint32_t key,
                                          const int* __restrict__ cstart) {
  const int nxy = K.nx * K.ny;
  const int cz = (int)key / nxy;
  const int rem = (int)key - cz * nxy;
  const int cy = rem / K.nx;
  const int cx = rem - cy * K.nx;
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx + 1 < K.nx ? cx + 1 : K.nx - 1;
  Rows9 r;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int rz = cz - 1 + k / 3, ry = cy - 1 + k % 3;
    if (rz < 0 || rz >= K.nz || ry < 0 || ry >= K.ny) {
      r.t0[k] = 0;
      r.t1[k] = 0;
    } else {
      const int base = (rz * K.ny + ry) * K.nx;
      r.t0[k] = cstart[base + x0];
      r.t1[k] = cstart[base + x1 + 1];
    }
  }
  return r;
}

// ------------------------------------------------------------------------
// pair arithmetic (the 2D forms with a z component)
// ------------------------------------------------------------------------
template <typename R, bool SUM>
__device__ __forceinline__ void dens3_pair(const KParams& K, R dx, R dy, R dz, R r2, R dvx,
                                           R dvy, R dvz, Acc<R>& sum, Acc<R>& rs) {
  const R q = fsqrt(r2) * (R)K.invh;
  const R tq = q - (R)2;
  const R F = (R)K.C1 * (tq * tq * tq);
  const R vcc = fmr(dvz, F * dz, fmr(dvy, F * dy, dvx * (F * dx)));
  sum.add((R)K.mass * vcc);
  if (SUM) {
    const R s = fmr((R)-0.5, q, (R)1);
    const R s2 = s * s;
    rs.add((R)K.mass * ((R)K.wnorm * (s2 * s2) * fmr((R)2, q, (R)1)));
  }
}

// Neighbour masks recorded by the density pass for the force pass: for row k
// of particle i, bit b of word q (q < M3_WORDS) says whether the candidate at
// sorted position t0_k + 32 q + b is a neighbour, t0_k the first position of
// the row's three cells.  Structure of arrays: word k * M3_WORDS + q of
// sorted particle i at w[(k * M3_WORDS + q) * stride + i]; word M3_OK is 1
// when every row had at most 32 M3_WORDS candidates (complete masks).
constexpr int M3_WORDS = 3;
constexpr int M3_OK = 9 * M3_WORDS;
constexpr int M3_TOTAL = M3_OK + 1;

struct Mask3 {
  uint32_t* w;
  int stride;
  __device__ __forceinline__ void put(int q, int i, uint32_t m) const {
    w[(size_t)q * stride + i] = m;
  }
  __device__ __forceinline__ uint32_t get(int q, int i) const {
    return w[(size_t)q * stride + i];
  }
};

template <typename R, bool SUM, bool EOS>
__device__ __forceinline__ void density3_direct(St3<R>& S, const KParams& K, int i,
                                                const uint32_t* __restrict__ skey,
                                                const int* __restrict__ cstart, Mask3 nbm) {
  const auto pi = load_pos3(S, i);
  const auto vi = S.v[i];
  const Rows9 rw = rows3_of(K, skey[i], cstart);
  const R cut2 = cutoff2<R>(K);
  DensOwn<R> o;
  dens_begin<R, SUM>(K, o);
  bool ok = true;
  // FP32 form: the neighbour pre-test of every row's candidates (at most 96
  // per row), eight at a time with no branch, into three mask words per row
  // kept in the lane's own LDS column; then one loop over the set bits of
  // all 27 words in (row, word, bit) order, each survivor through the exact
  // test and the pair sums (a wavefront runs for its longest lane's survivor
  // count, not for the sum of the nine per-row maxima).  A lane with a row of
  // more than 96 candidates tests and sums its candidates one by one
  // instead.  Either way the pairs are summed in ascending position, row by
  // row, as the single test-and-sum walk of the FP64 form (below) does.
  if constexpr (sizeof(R) == 4) {
    constexpr int NB = 8;
    __shared__ uint32_t lw[M3_OK][BLOCK];
    __shared__ int lt[9][BLOCK];
    const int lane = threadIdx.x;
    // pre-test positions: x = cx cell + off rounded once (K.pre2f bounds
    // the exact test's cut-off through that rounding)
    const float cellf = K.cellf;
    const float xi = __builtin_fmaf(pi.cxf, cellf, pi.ox);
    const float pre2 = K.pre2f;
    uint32_t nz = 0u;  // the lane's non-empty words
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int t0 = rw.t0[k], t1 = rw.t1[k];
      if (t1 - t0 > 32 * M3_WORDS) ok = false;
      const int tA = t1 < t0 + 32 * M3_WORDS ? t1 : t0 + 32 * M3_WORDS;
      const float yi = pi.oy + row_u3<float>(K, 1 - k % 3), zi = pi.oz + row_u3<float>(K, 1 - k / 3);
      uint32_t m0 = 0, m1 = 0, m2 = 0;
      for (int c0 = t0; c0 < tA; c0 += NB) {
        // a batch starts a multiple of NB past t0, so its bits lie in one
        // mask word: test into a byte, then place the byte once.  The
        // position records are padded by NB - 1 past the last particle.
        uint32_t bb = 0u;
        const float4* pb = S.pf + c0;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const float4 pj = pb[q];
          const float dx = xi - __builtin_fmaf(pj.w, cellf, pj.x);
          const float dy = yi - pj.y;
          const float dz = zi - pj.z;
          const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
          bb |= (r2 < pre2 ? 1u : 0u) << q;
        }
        const int len = tA - c0;
        if (len < NB) bb &= (1u << len) - 1u;
        const int b0 = c0 - t0;
        const uint32_t mb = bb << (b0 & 31);
        m0 |= b0 < 32 ? mb : 0u;
        m1 |= (b0 >= 32 && b0 < 64) ? mb : 0u;
        m2 |= b0 >= 64 ? mb : 0u;
      }
      {
        // the lane's own position passed the test at r = 0: clear it
        const int sb = i - t0;
        if (sb >= 0 && sb < tA - t0) {
          const uint32_t keep = ~(1u << (sb & 31));
          m0 &= sb < 32 ? keep : ~0u;
          m1 &= (sb >= 32 && sb < 64) ? keep : ~0u;
          m2 &= sb >= 64 ? keep : ~0u;
        }
      }
      lw[3 * k + 0][lane] = m0;
      lw[3 * k + 1][lane] = m1;
      lw[3 * k + 2][lane] = m2;
      nz |= (m0 != 0u ? 1u : 0u) << (3 * k);
      nz |= (m1 != 0u ? 2u : 0u) << (3 * k);
      nz |= (m2 != 0u ? 4u : 0u) << (3 * k);
      lt[k][lane] = t0;
    }
    if (ok) {
      // the word's first candidate and its row's y and z terms are found
      // when the walk enters the word; a survivor that fails the exact test
      // leaves the lane's copy of the word
      uint32_t m = 0u;
      int tb = 0, wc = 0;
      float uy = 0.f, uz = 0.f;
      while (true) {
        if (m == 0u) {
          if (nz == 0u) break;
          const int w = __builtin_ctz(nz);
          nz &= nz - 1u;
          m = lw[w][lane];
          wc = w;
          const int k = (w * 11) >> 5;  // w / 3 for w < 27
          const int q = w - 3 * k;
          const int ky = (k * 11) >> 5;  // k / 3 for k < 9
          uy = row_u3<float>(K, 1 - (k - 3 * ky));
          uz = row_u3<float>(K, 1 - ky);
          tb = lt[k][lane] + 32 * q;
        }
        const int b = __builtin_ctz(m);
        m &= m - 1u;
        const int t = tb + b;
        const auto pj = load_pos3(S, t);
        const auto vj = S.v[t];
        float dx, dy, dz;
        const float r2 = delta3(pi, pj, K, uy, uz, dx, dy, dz);
        // both loads in flight together: keep the velocity load out of
        // the branch below
        asm volatile("" ::"v"(vj.x), "v"(vj.y), "v"(vj.z));
        if (!(r2 < cut2)) {
          lw[wc][lane] &= ~(1u << b);  // passed the pre-test only
          continue;
        }
        dens3_pair<R, SUM>(K, dx, dy, dz, r2, vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, o.sum,
                           o.rs);
      }
      if (nbm.w)
        for (int q = 0; q < M3_OK; ++q) nbm.put(q, i, lw[q][lane]);
    } else {
      // a row of more than 96 candidates: test and sum one by one, in the
      // same order; masks of the first 96 candidates of each row
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const int t0 = rw.t0[k], t1 = rw.t1[k];
        const float uy = row_u3<float>(K, 1 - k % 3), uz = row_u3<float>(K, 1 - k / 3);
        uint32_t m0 = 0, m1 = 0, m2 = 0;
        for (int t = t0; t < t1; ++t) {
          if (t == i) continue;
          const auto pj = load_pos3(S, t);
          float dx, dy, dz;
          const float r2 = delta3(pi, pj, K, uy, uz, dx, dy, dz);
          if (!(r2 < cut2)) continue;
          const int b = t - t0;
          const uint32_t bit = 1u << (b & 31);
          m0 |= b < 32 ? bit : 0u;
          m1 |= (b >= 32 && b < 64) ? bit : 0u;
          m2 |= (b >= 64 && b < 96) ? bit : 0u;
          const auto vj = S.v[t];
          dens3_pair<R, SUM>(K, dx, dy, dz, r2, vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, o.sum,
                             o.rs);
        }
        if (nbm.w) {
          nbm.put(k * M3_WORDS + 0, i, m0);
          nbm.put(k * M3_WORDS + 1, i, m1);
          nbm.put(k * M3_WORDS + 2, i, m2);
        }
      }
    }
  } else {
    for (int k = 0; k < 9; ++k) {
      const int t0 = rw.t0[k], t1 = rw.t1[k];
      const R uy = row_u3<R>(K, 1 - k % 3), uz = row_u3<R>(K, 1 - k / 3);
      if (t1 - t0 > 32 * M3_WORDS) ok = false;
      uint32_t m0 = 0, m1 = 0, m2 = 0;
      for (int t = t0; t < t1; ++t) {
        if (t == i) continue;
        const auto pj = load_pos3(S, t);
        R dx, dy, dz;
        const R r2 = delta3(pi, pj, K, uy, uz, dx, dy, dz);
        if (!(r2 < cut2)) continue;
        const int b = t - t0;
        const uint32_t bit = 1u << (b & 31);
        m0 |= b < 32 ? bit : 0u;
        m1 |= (b >= 32 && b < 64) ? bit : 0u;
        m2 |= (b >= 64 && b < 96) ? bit : 0u;
        const auto vj = S.v[t];
        dens3_pair<R, SUM>(K, dx, dy, dz, r2, vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, o.sum, o.rs);
      }
      if (nbm.w) {
        nbm.put(k * M3_WORDS + 0, i, m0);
        nbm.put(k * M3_WORDS + 1, i, m1);
        nbm.put(k * M3_WORDS + 2, i, m2);
      }
    }
  }
  if (nbm.w) nbm.put(M3_OK, i, ok ? 1u : 0u);
  dens_end<R, SUM, EOS>(S, K, i, o);
}

template <typename R, bool SUM, bool EOS>
__global__ __launch_bounds__(BLOCK) void k3_density(St3<R> S, KParams K,
                                                    const uint32_t* __restrict__ skey,
                                                    const int* __restrict__ cstart, Mask3 nbm) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  density3_direct<R, SUM, EOS>(S, K, i, skey, cstart, nbm);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_eos(St3<R> S, KParams K) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  eos_point(S, K, i, is_rate(K, S.id[i]), S.rho[i], S.divv[i]);
}

// Force pass: int_force pressure part, art_visc, art_heat and av_vel in the
// Q = 1/rho, A = p/rho form of the 2D force_pair (same terms per pair).
template <typename R> struct ForceOwn3 {
  using R4 = typename V4<R>::type;
  R4 v;
  R Q, A, psi, u;
  R fx = 0, fy = 0, fz = 0, xsx = 0, xsy = 0, xsz = 0;
  Acc<R> du;
};

template <typename R>
__device__ __forceinline__ void force3_begin(const St3<R>& S, const KParams& K, int i,
                                             ForceOwn3<R>& o) {
  o.v = S.v[i];
  q_and_a(K, S.rho[i], S.p[i], o.Q, o.A);
  o.psi = S.psi[i];
  if constexpr (sizeof(R) == 4) o.u = S.uf[i]; else o.u = (R)S.u[i];
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS>
__device__ __forceinline__ void force3_pair(const KParams& K, ForceOwn3<R>& o, R dx, R dy, R dz,
                                            R r2, R vjx, R vjy, R vjz, R Qj, R Aj, R psij,
                                            R uj) {
  const R m = (R)K.mass;
  const R q = fsqrt(r2) * (R)K.invh;
  const R tq = q - (R)2;
  const R F = (R)K.C1 * (tq * tq * tq);
  const R dvx = o.v.x - vjx, dvy = o.v.y - vjy, dvz = o.v.z - vjz;
  const R vr = fmr(dvz, dz, fmr(dvy, dy, dvx * dx));
  const R inv_mrho = (R)2 * o.Q * Qj * frcp(o.Q + Qj);
  R coef = 0;
  if (PRESS) coef = (K.pa_sph == 0) ? fmr(o.A, Qj, Aj * o.Q) : fmr(o.A, o.Q, Aj * Qj);
  R inv_ea = 0;
  if (AV) {
    inv_ea = frcp(r2 + (R)K.av_e2);
    if (vr < (R)0) {
      const R phi = (R)K.h * vr * inv_ea;
      const R pc = fmr((R)K.av_beta, phi, -((R)K.av_alpha * (R)K.c0));
      coef = fmr(pc * phi, inv_mrho, coef);
    }
  }
  if (PRESS || AV) {
    const R G = m * F * coef;
    o.fx = fmr(-G, dx, o.fx);
    o.fy = fmr(-G, dy, o.fy);
    o.fz = fmr(-G, dz, o.fz);
    R dut = (R)0.5 * G * vr;
    if (AH) {
      const R inv_eh = (AV && K.same_e2) ? inv_ea : frcp(r2 + (R)K.ah_e2);
      dut = fmr(m * (o.psi + psij) * F * r2 * inv_mrho * inv_eh, o.u - uj, dut);
    }
    o.du.add(dut);
  } else if (AH) {
    const R inv_eh = frcp(r2 + (R)K.ah_e2);
    o.du.add(m * (o.psi + psij) * F * r2 * inv_mrho * inv_eh * (o.u - uj));
  }
  if (XS) {
    const R s = fmr((R)-0.5, q, (R)1);
    const R s2 = s * s;
    const R w = (R)K.wnorm * (s2 * s2) * fmr((R)2, q, (R)1);
    const R wf = (R)K.eps_m * inv_mrho * w;
    o.xsx = fmr(-wf, dvx, o.xsx);
    o.xsy = fmr(-wf, dvy, o.xsy);
    o.xsz = fmr(-wf, dvz, o.xsz);
  }
}

// The deposition heat source after the 3D force pass (fluka_heat.f:235-337),
// as k_deposition in 2D.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_deposition(St3<R> S, KParams K) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n || S.id[i] >= K.np) return;
  const double4 q = S.xyz[i];
  R Q, A;
  q_and_a(K, S.rho[i], S.p[i], Q, A);
  S.du[i] = S.du[i] + (R)deposition_at(K, q.x, q.y, q.z) * Q;
}

template <typename R, bool PRESS, bool XS>
__device__ __forceinline__ void force3_end(St3<R>& S, const KParams& K, int i,
                                           const ForceOwn3<R>& o) {
  using R4 = typename V4<R>::type;
  if (PRESS) {
    const int id = S.id[i];
    const bool grav = (K.flags & 16) && id < K.np;
    R4 acc;
    acc.x = o.fx + (grav ? (R)K.gx : (R)0);
    acc.y = o.fy + (grav ? (R)K.gy : (R)0);
    acc.z = o.fz + (grav ? (R)K.gz : (R)0);
    acc.w = (R)0;
    S.a[i] = acc;
    S.du[i] = o.du.get();
  }
  if (XS) {
    R4 w;
    w.x = o.xsx;
    w.y = o.xsy;
    w.z = o.xsz;
    w.w = (R)0;
    S.av[i] = w;
  }
}

template <typename R> __device__ __forceinline__ R u3_of(const St3<R>& S, int t) {
  if constexpr (sizeof(R) == 4) return S.uf[t]; else return (R)S.u[t];
}

// Neighbour data of the force pass packed per particle into one record of
// three vectors, so that a neighbour's loads fall in one or two cache lines:
// the position record of the state (FP32: pf; FP64: xyz), (v, Q = 1/rho) and
// (A = p/rho, psi, u, 0), written by k3_force_pack right before the gather
// with q_and_a's arithmetic.
template <typename R> struct Pack3 {
  using R4 = typename V4<R>::type;
  R4* rec;  // 3 per particle
};

__device__ __forceinline__ float4 pos_rec3(const St3<float>& S, int i) { return S.pf[i]; }
__device__ __forceinline__ double4 pos_rec3(const St3<double>& S, int i) { return S.xyz[i]; }
__device__ __forceinline__ Pos3F pos_of_rec(float4 q) { return {q.w, q.x, q.y, q.z}; }
__device__ __forceinline__ Pos3D pos_of_rec(double4 q) { return {q.x, q.y, q.z}; }

template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_force_pack(St3<R> S, KParams K, Pack3<R> P) {
  if (halted(K)) return;
  using R4 = typename V4<R>::type;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  R Q, A;
  q_and_a(K, S.rho[i], S.p[i], Q, A);
  const R4 v = S.v[i];
  R4 vq, ap;
  vq.x = v.x;
  vq.y = v.y;
  vq.z = v.z;
  vq.w = Q;
  ap.x = A;
  ap.y = S.psi[i];
  ap.z = u3_of(S, i);
  ap.w = (R)0;
  P.rec[3 * (size_t)i] = pos_rec3(S, i);
  P.rec[3 * (size_t)i + 1] = vq;
  P.rec[3 * (size_t)i + 2] = ap;
}

// One neighbour t of particle i in the direct path.
template <typename R, bool PRESS, bool AV, bool AH, bool XS, typename P3>
__device__ __forceinline__ void force3_visit(const St3<R>& S, const Pack3<R>& P, const KParams& K,
                                             const P3& pi, R uy, R uz, ForceOwn3<R>& o, int t) {
  const auto* rp = P.rec + (((unsigned)t << 1) + (unsigned)t);  // 3 t, no 64-bit multiply
  const auto pj = pos_of_rec(rp[0]);
  const auto vq = rp[1];
  const auto ap = rp[2];
  R dx, dy, dz;
  const R r2 = delta3(pi, pj, K, uy, uz, dx, dy, dz);
  force3_pair<R, PRESS, AV, AH, XS>(K, o, dx, dy, dz, r2, vq.x, vq.y, vq.z, vq.w, ap.x, ap.y,
                                    ap.z);
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS>
__device__ __forceinline__ void force3_direct(St3<R>& S, const Pack3<R>& P, const KParams& K,
                                              int i, const uint32_t* __restrict__ skey,
                                              const int* __restrict__ cstart, Mask3 nbm) {
  const auto pi = load_pos3(S, i);
  ForceOwn3<R> o;
  force3_begin(S, K, i, o);
  const Rows9 rw = rows3_of(K, skey[i], cstart);
  if (nbm.w && nbm.get(M3_OK, i)) {
    // One loop over every set bit of the 27 words in (row, word, bit) order,
    // the order of the per-word walk: a wavefront runs for its longest
    // lane's neighbour count instead of the sum of the 27 per-word maxima.
    // Each lane keeps its words and row starts in its own LDS column (they
    // are indexed by the lane's position in the walk) and a bitmap of its
    // non-empty words.
    __shared__ uint32_t lw[M3_OK][BLOCK];
    __shared__ int lt[9][BLOCK];
    const int lane = threadIdx.x;
    uint32_t nz = 0u;
#pragma unroll
    for (int q = 0; q < M3_OK; ++q) {
      const uint32_t m = nbm.get(q, i);
      lw[q][lane] = m;
      nz |= (m != 0u ? 1u : 0u) << q;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) lt[k][lane] = rw.t0[k];
    if (nz) {
      // the word's first candidate and its row's y and z terms are found
      // when the walk enters the word, and kept until it leaves it
      uint32_t m = 0u;
      int tb = 0;
      R uy = (R)0, uz = (R)0;
      while (true) {
        if (m == 0u) {
          if (nz == 0u) break;
          const int w = __builtin_ctz(nz);
          nz &= nz - 1u;
          m = lw[w][lane];
          const int k = (w * 11) >> 5;  // w / 3 for w < 27
          const int q = w - 3 * k;
          const int ky = (k * 11) >> 5;  // k / 3 for k < 9
          uy = row_u3<R>(K, 1 - (k - 3 * ky));
          uz = row_u3<R>(K, 1 - ky);
          tb = lt[k][lane] + 32 * q;
        }
        const int b = __builtin_ctz(m);
        m &= m - 1u;
        force3_visit<R, PRESS, AV, AH, XS>(S, P, K, pi, uy, uz, o, tb + b);
      }
    }
  } else {
    const R cut2 = cutoff2<R>(K);
    for (int k = 0; k < 9; ++k) {
      const R uy = row_u3<R>(K, 1 - k % 3), uz = row_u3<R>(K, 1 - k / 3);
      for (int t = rw.t0[k]; t < rw.t1[k]; ++t) {
        if (t == i) continue;
        const auto pj = load_pos3(S, t);
        R dx, dy, dz;
        const R r2 = delta3(pi, pj, K, uy, uz, dx, dy, dz);
        if (!(r2 < cut2)) continue;
        force3_visit<R, PRESS, AV, AH, XS>(S, P, K, pi, uy, uz, o, t);
      }
    }
  }
  force3_end<R, PRESS, XS>(S, K, i, o);
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS>
__global__ __launch_bounds__(BLOCK) void k3_force(St3<R> S, Pack3<R> P, KParams K,
                                                  const uint32_t* __restrict__ skey,
                                                  const int* __restrict__ cstart, Mask3 nbm) {
  if (halted(K)) return;
  // contiguous shares per XCD: the packed records of the neighbour rows
  // stay in one L2 (3.6x less traffic beyond L2, 915 -> 880 us in FP32);
  // the density pass, bound by instruction issue, keeps the plain order,
  // where the XCDs' shares of work are even
  const int i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  force3_direct<R, PRESS, AV, AH, XS>(S, P, K, i, skey, cstart, nbm);
}

// ------------------------------------------------------------------------
// leapfrog (time_integration.f:136-154 predictor, :186-217 corrector)
// ------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_predictor(St3<R> S, KParams K) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const int id = S.id[i];
  const double hdt = K.dt / 2.0;
  S.flag[i] = 0;
  if (is_rate(K, id)) {
    double u = S.u[i];
    S.u_min[i] = u;
    u = u + hdt * (double)S.du[i];
    if (u < 0.0) {
      u = 0.0;
      S.flag[i] = 1u;
    }
    S.u[i] = u;
    if constexpr (sizeof(R) == 4) S.uf[i] = (float)u;
    if (K.density_method == 0) {
      const R rho = S.rho[i];
      S.rho_min[i] = rho;
      S.rho[i] = rho + (R)hdt * S.drho[i];
    }
  }
  if (id < K.np) {
    auto v = S.v[i];
    const auto a = S.a[i];
    S.v_min[i] = v;
    v.x = v.x + (R)hdt * a.x;
    v.y = v.y + (R)hdt * a.y;
    v.z = v.z + (R)hdt * a.z;
    S.v[i] = v;
  }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_corrector(St3<R> S, KParams K, int its, int keys,
                                                      KeyOut ko, int* __restrict__ err) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = i < K.n;
  uint32_t key = 0;
  int id = 0;
  if (valid) {
    id = S.id[i];
    const double dt = K.dt, hdt = K.dt / 2.0;
    double4 q = S.xyz[i];
    if (is_rate(K, id)) {
      double u;
      if (its != 1) {
        u = S.u_min[i] + dt * (double)S.du[i];
        if (K.density_method == 0) S.rho[i] = S.rho_min[i] + (R)dt * S.drho[i];
      } else {
        u = S.u[i] + hdt * (double)S.du[i];
        if (K.density_method == 0) S.rho[i] = S.rho[i] + (R)hdt * S.drho[i];
      }
      if (u < 0.0) u = 0.0;
      S.u[i] = u;
      if constexpr (sizeof(R) == 4) S.uf[i] = (float)u;
      if (id < K.np) {
        auto v = S.v[i];
        const auto a = S.a[i];
        if (its != 1) {
          const auto vm = S.v_min[i];
          const auto av = S.av[i];
          v.x = vm.x + (R)dt * a.x + av.x;
          v.y = vm.y + (R)dt * a.y + av.y;
          v.z = vm.z + (R)dt * a.z + av.z;
        } else {
          v.x = v.x + (R)hdt * a.x;
          v.y = v.y + (R)hdt * a.y;
          v.z = v.z + (R)hdt * a.z;
        }
        S.v[i] = v;
        q.x = q.x + dt * (double)v.x;
        q.y = q.y + dt * (double)v.y;
        q.z = q.z + dt * (double)v.z;
        S.xyz[i] = q;
      }
    }
    if (keys) {
      int ix, iy, iz;
      bool ok;
      key = cell_of3(K, q.x, q.y, q.z, ix, iy, iz, ok);
      if (!ok) {
      atomicAdd(err, 1);
      raise_halt(K);
    }
    }
  }
  if (keys) emit_key(ko, valid, i, id, key);
}

// ------------------------------------------------------------------------
// pair census: neighbours with larger id
// ------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_pair_count(St3<R> S, KParams K,
                                                       const uint32_t* __restrict__ skey,
                                                       const int* __restrict__ cstart,
                                                       int* __restrict__ cnt,
                                                       const int* __restrict__ offs,
                                                       int* __restrict__ out_i,
                                                       int* __restrict__ out_j) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const auto pi = load_pos3(S, i);
  const int idi = S.id[i];
  const Rows9 rw = rows3_of(K, skey[i], cstart);
  const R cut2 = cutoff2<R>(K);
  int c = 0;
  const int o = offs ? offs[i] : 0;
  for (int k = 0; k < 9; ++k) {
    const R uy = row_u3<R>(K, 1 - k % 3), uz = row_u3<R>(K, 1 - k / 3);
    for (int t = rw.t0[k]; t < rw.t1[k]; ++t) {
      const int idj = S.id[t];
      if (idj <= idi) continue;
      const auto pj = load_pos3(S, t);
      R dx, dy, dz;
      const R r2 = delta3(pi, pj, K, uy, uz, dx, dy, dz);
      if (!(r2 < cut2)) continue;
      if (offs) {
        out_i[o + c] = idi;
        out_j[o + c] = idj;
      }
      ++c;
    }
  }
  if (cnt) cnt[i] = c;
}

// FP64 sums over real particles: mass, momentum x, y, z, kinetic and
// internal energy; one row of partial sums per block.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_conserve(St3<R> S, KParams K,
                                                     double* __restrict__ part) {
  __shared__ double sh[6][BLOCK];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < K.n && S.id[i] < K.np) {
    const auto vel = S.v[i];
    const double vx = (double)vel.x, vy = (double)vel.y, vz = (double)vel.z;
    v[0] = K.mass;
    v[1] = K.mass * vx;
    v[2] = K.mass * vy;
    v[3] = K.mass * vz;
    v[4] = 0.5 * K.mass * (vx * vx + vy * vy + vz * vz);
    v[5] = K.mass * S.u[i];
  }
  for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) part[blockIdx.x * 6 + k] = sh[k][0];
}

}  // namespace arm
