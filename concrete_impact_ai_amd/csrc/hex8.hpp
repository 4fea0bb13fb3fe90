// Hex8 kinematics on uniform box elements (gfx950 device helpers).
//
// Shape functions follow src/fem/elements/lagrange.py:50-76 (local node
// a = ax + 2 ay + 4 az, x fastest) and the 2x2x2 Gauss rule of
// src/fem/quadrature/rules.py:24-33 (point q = 4 qx + 2 qy + qz, zeta fastest,
// coordinates -1/sqrt(3), +1/sqrt(3), unit weights).  For a box element of
// edges (hx, hy, hz) the Jacobian is constant, so
//     dN_a/dx (q) = s(ax)/hx * W(qy, ay) * W(qz, az),   W(q, a) = (1 + s(a) p(q)) / 2,
// and w_q det J = hx hy hz / 8.  The 6x24 B matrix of
// src/fem/assembly/elasticity.py:160-183 is applied without being formed.
#pragma once

#include <hip/hip_runtime.h>

namespace cia {

constexpr double kGp = 0.57735026918962576451;  // 1/sqrt(3)
constexpr double kWp = 0.5 * (1.0 + kGp);
constexpr double kWm = 0.5 * (1.0 - kGp);

__host__ __device__ constexpr double wq(int q, int a) { return q == a ? kWp : kWm; }
__host__ __device__ constexpr double sgn(int a) { return a ? 1.0 : -1.0; }

// Reference-coordinate gradient factors (without the 1/h scaling).
__host__ __device__ constexpr double dnx(int q, int a) {
  return sgn(a & 1) * wq((q >> 1) & 1, (a >> 1) & 1) * wq(q & 1, (a >> 2) & 1);
}
__host__ __device__ constexpr double dny(int q, int a) {
  return sgn((a >> 1) & 1) * wq((q >> 2) & 1, a & 1) * wq(q & 1, (a >> 2) & 1);
}
__host__ __device__ constexpr double dnz(int q, int a) {
  return sgn((a >> 2) & 1) * wq((q >> 2) & 1, a & 1) * wq((q >> 1) & 1, (a >> 1) & 1);
}

struct BoxGeom {
  double ihx, ihy, ihz;  // 1 / element edge
  double jw;             // w_q det J
};

// Engineering strain at Gauss point Q from nodal values u[a][c].
template <int Q, class T>
__device__ __forceinline__ void strain_at(const T u[8][3], const BoxGeom& g, T e[6]) {
  T gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gx[c] += (T)dnx(Q, a) * u[a][c];
      gy[c] += (T)dny(Q, a) * u[a][c];
      gz[c] += (T)dnz(Q, a) * u[a][c];
    }
  }
  const T ix = (T)g.ihx, iy = (T)g.ihy, iz = (T)g.ihz;
  e[0] = ix * gx[0];
  e[1] = iy * gy[1];
  e[2] = iz * gz[2];
  e[3] = iz * gz[1] + iy * gy[2];
  e[4] = iz * gz[0] + ix * gx[2];
  e[5] = iy * gy[0] + ix * gx[1];
}

// f[a][c] += w detJ (B_Q^T sigma)[a][c].
template <int Q, class T>
__device__ __forceinline__ void force_at(const T s[6], const BoxGeom& g, T f[8][3]) {
  const T jx = (T)(g.jw * g.ihx), jy = (T)(g.jw * g.ihy), jz = (T)(g.jw * g.ihz);
  const T tx0 = jx * s[0], tx1 = jx * s[5], tx2 = jx * s[4];
  const T ty0 = jy * s[5], ty1 = jy * s[1], ty2 = jy * s[3];
  const T tz0 = jz * s[4], tz1 = jz * s[3], tz2 = jz * s[2];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const T bx = (T)dnx(Q, a), by = (T)dny(Q, a), bz = (T)dnz(Q, a);
    f[a][0] += bx * tx0 + by * ty0 + bz * tz0;
    f[a][1] += bx * tx1 + by * ty1 + bz * tz1;
    f[a][2] += bx * tx2 + by * ty2 + bz * tz2;
  }
}

// Strain columns of B_Q for node a (unit displacement in x, y, z).
template <class T>
__host__ __device__ __forceinline__ void b_columns(int q, int a, const BoxGeom& g, T col[3][6]) {
  const T bx = (T)(dnx(q, a) * g.ihx), by = (T)(dny(q, a) * g.ihy), bz = (T)(dnz(q, a) * g.ihz);
  // unit u_x: (bx, 0, 0, 0, bz, by); u_y: (0, by, 0, bz, 0, bx); u_z: (0, 0, bz, by, bx, 0)
  col[0][0] = bx; col[0][1] = 0;  col[0][2] = 0;  col[0][3] = 0;  col[0][4] = bz; col[0][5] = by;
  col[1][0] = 0;  col[1][1] = by; col[1][2] = 0;  col[1][3] = bz; col[1][4] = 0;  col[1][5] = bx;
  col[2][0] = 0;  col[2][1] = 0;  col[2][2] = bz; col[2][3] = by; col[2][4] = bx; col[2][5] = 0;
}

// Gradient factors of the 8 corner shape functions at a runtime Gauss point q
// (already scaled by 1/h): dN_a/dx = bx[a], dN_a/dy = by[a], dN_a/dz = bz[a].
template <class T>
__device__ __forceinline__ void gp_factors(int q, const BoxGeom& g, T bx[8], T by[8], T bz[8]) {
  const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
  const T wx[2] = {(T)(qx == 0 ? kWp : kWm), (T)(qx == 1 ? kWp : kWm)};
  const T wy[2] = {(T)(qy == 0 ? kWp : kWm), (T)(qy == 1 ? kWp : kWm)};
  const T wz[2] = {(T)(qz == 0 ? kWp : kWm), (T)(qz == 1 ? kWp : kWm)};
  const T ix = (T)g.ihx, iy = (T)g.ihy, iz = (T)g.ihz;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
    bx[a] = (ax ? ix : -ix) * wy[ay] * wz[az];
    by[a] = (ay ? iy : -iy) * wx[ax] * wz[az];
    bz[a] = (az ? iz : -iz) * wx[ax] * wy[ay];
  }
}

// Gradient factors of corner a at the Gauss point with per-direction weights
// w?[0|1] (scaled by 1/h via ih?).
template <class T>
__device__ __forceinline__ void corner_factors(int a, const T wx[2], const T wy[2], const T wz[2],
                                               T ix, T iy, T iz, T& bx, T& by, T& bz) {
  const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
  bx = (ax ? ix : -ix) * wy[ay] * wz[az];
  by = (ay ? iy : -iy) * wx[ax] * wz[az];
  bz = (az ? iz : -iz) * wx[ax] * wy[ay];
}

template <class T>
__device__ __forceinline__ void gp_weights(int q, T wx[2], T wy[2], T wz[2]) {
  const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
  wx[0] = (T)(qx == 0 ? kWp : kWm);
  wx[1] = (T)(qx == 1 ? kWp : kWm);
  wy[0] = (T)(qy == 0 ? kWp : kWm);
  wy[1] = (T)(qy == 1 ? kWp : kWm);
  wz[0] = (T)(qz == 0 ? kWp : kWm);
  wz[1] = (T)(qz == 1 ? kWp : kWm);
}

// Engineering strains of two nodal fields at once, read corner by corner from
// rows[a][0..2] (field 1) and rows[a][off..off+2] (field 2), so the 48 corner values
// are never held in registers together.
template <class T, int STRIDE>
__device__ __forceinline__ void strain_pair_from_rows(const double (*rows)[STRIDE], int off, int q,
                                                      const BoxGeom& geo, T e1[6], T e2[6]) {
  T wx[2], wy[2], wz[2];
  gp_weights<T>(q, wx, wy, wz);
  const T ix = (T)geo.ihx, iy = (T)geo.ihy, iz = (T)geo.ihz;
  T g1[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, g2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    T bx, by, bz;
    corner_factors<T>(a, wx, wy, wz, ix, iy, iz, bx, by, bz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v1 = (T)rows[a][c], v2 = (T)rows[a][off + c];
      g1[c][0] += bx * v1;
      g1[c][1] += by * v1;
      g1[c][2] += bz * v1;
      g2[c][0] += bx * v2;
      g2[c][1] += by * v2;
      g2[c][2] += bz * v2;
    }
  }
  e1[0] = g1[0][0];
  e1[1] = g1[1][1];
  e1[2] = g1[2][2];
  e1[3] = g1[1][2] + g1[2][1];
  e1[4] = g1[0][2] + g1[2][0];
  e1[5] = g1[0][1] + g1[1][0];
  e2[0] = g2[0][0];
  e2[1] = g2[1][1];
  e2[2] = g2[2][2];
  e2[3] = g2[1][2] + g2[2][1];
  e2[4] = g2[0][2] + g2[2][0];
  e2[5] = g2[0][1] + g2[1][0];
}

// f[a][c] = w detJ (B^T sigma)[a][c] at Gauss point q, factors formed per corner.
template <class T>
__device__ __forceinline__ void force_at_gp(const T s[6], T jw, int q, const BoxGeom& geo,
                                            T f[8][3]) {
  T wx[2], wy[2], wz[2];
  gp_weights<T>(q, wx, wy, wz);
  const T ix = (T)geo.ihx, iy = (T)geo.ihy, iz = (T)geo.ihz;
  const T s0 = jw * s[0], s1 = jw * s[1], s2 = jw * s[2], s3 = jw * s[3], s4 = jw * s[4],
          s5 = jw * s[5];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    T bx, by, bz;
    corner_factors<T>(a, wx, wy, wz, ix, iy, iz, bx, by, bz);
    f[a][0] = bx * s0 + by * s5 + bz * s4;
    f[a][1] = by * s1 + bz * s3 + bx * s5;
    f[a][2] = bz * s2 + by * s3 + bx * s4;
  }
}

// Engineering strain at the point whose factors are given.
template <class T, class U>
__device__ __forceinline__ void strain_from(const U u[8][3], const T bx[8], const T by[8],
                                            const T bz[8], T e[6]) {
  T g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // g[c][d] = du_c / dx_d
#pragma unroll
  for (int a = 0; a < 8; ++a) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = (T)u[a][c];
      g[c][0] += bx[a] * v;
      g[c][1] += by[a] * v;
      g[c][2] += bz[a] * v;
    }
  }
  e[0] = g[0][0];
  e[1] = g[1][1];
  e[2] = g[2][2];
  e[3] = g[1][2] + g[2][1];
  e[4] = g[0][2] + g[2][0];
  e[5] = g[0][1] + g[1][0];
}

// f[a][c] = w detJ (B^T sigma)[a][c] at one point.
template <class T>
__device__ __forceinline__ void force_from(const T s[6], T jw, const T bx[8], const T by[8],
                                           const T bz[8], T f[8][3]) {
  const T s0 = jw * s[0], s1 = jw * s[1], s2 = jw * s[2], s3 = jw * s[3], s4 = jw * s[4],
          s5 = jw * s[5];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    f[a][0] = bx[a] * s0 + by[a] * s5 + bz[a] * s4;
    f[a][1] = by[a] * s1 + bz[a] * s3 + bx[a] * s5;
    f[a][2] = bz[a] * s2 + by[a] * s3 + bx[a] * s4;
  }
}

// Value of the lane whose index differs in bit MASK (1, 2 or 4) inside an
// aligned group of 8 lanes, by DPP moves on the vector ALU (quad permutes for
// bits 0 and 1, row shifts by 4 for bit 2) instead of LDS-crossbar shuffles.  q
// is the lane index mod 8.  All lanes of the group must be active.
template <int MASK>
__device__ __forceinline__ int lane_xor(int v, int q) {
  if (MASK == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // [1,0,3,2]
  if (MASK == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // [2,3,0,1]
  const int from_up = __builtin_amdgcn_mov_dpp(v, 0x104, 0xF, 0xF, true);  // lane + 4
  const int from_dn = __builtin_amdgcn_mov_dpp(v, 0x114, 0xF, 0xF, true);  // lane - 4
  return (q & 4) ? from_dn : from_up;
}

template <int MASK>
__device__ __forceinline__ float lane_xor(float v, int q) {
  return __int_as_float(lane_xor<MASK>(__float_as_int(v), q));
}

template <int MASK>
__device__ __forceinline__ double lane_xor(double v, int q) {
  const int lo = lane_xor<MASK>(__double2loint(v), q);
  const int hi = lane_xor<MASK>(__double2hiint(v), q);
  return __hiloint2double(hi, lo);
}

// Sum f over the 8 lanes of an element (lanes q = 0..7 of an aligned 8-lane
// group, one per Gauss point) by recursive halving: 12 + 6 + 3 lane exchanges,
// after which lane q holds the element force at its corner q.  Fixed order, so
// the result is bitwise reproducible.
template <class T>
__device__ __forceinline__ void element_reduce8(int q, const T f[8][3], T out[3]) {
  T h1[4][3];
  const bool b4 = (q & 4) != 0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const T send = b4 ? f[j][m] : f[4 + j][m];
      const T recv = lane_xor<4>(send, q);
      h1[j][m] = (b4 ? f[4 + j][m] : f[j][m]) + recv;
    }
  T h2[2][3];
  const bool b2 = (q & 2) != 0;
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const T send = b2 ? h1[j][m] : h1[2 + j][m];
      const T recv = lane_xor<2>(send, q);
      h2[j][m] = (b2 ? h1[2 + j][m] : h1[j][m]) + recv;
    }
  const bool b1 = (q & 1) != 0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const T send = b1 ? h2[0][m] : h2[1][m];
    const T recv = lane_xor<1>(send, q);
    out[m] = (b1 ? h2[1][m] : h2[0][m]) + recv;
  }
}

// Eight consecutive values (one element's Gauss points) in 16-byte loads.
__device__ __forceinline__ void load8(const float* p, float v[8]) {
  const float4 a = *(const float4*)p, b = *(const float4*)(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void load8(const double* p, double v[8]) {
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const double2 a = *(const double2*)(p + 2 * h);
    v[2 * h] = a.x;
    v[2 * h + 1] = a.y;
  }
}

// Sum-factorised kinematics of one box element at all 8 Gauss points (one lane
// per element).  On a box element du/dx does not depend on the point's x
// coordinate: the x-gradients at the 4 (qy, qz) points are the 4 edge
// differences along x interpolated in y and then z, and likewise for y and z
// -- 180 operations for the 72 gradient entries instead of 8 x 72 products.
// gX[qy][qz][c] = du_c/dx, gY[qx][qz][c] = du_c/dy, gZ[qx][qy][c] = du_c/dz.
template <class T>
__device__ __forceinline__ void box_gradients(const T u[8][3], const BoxGeom& g, T gX[2][2][3],
                                              T gY[2][2][3], T gZ[2][2][3]) {
  const T wp = (T)kWp, wm = (T)kWm;
  const T ix = (T)g.ihx, iy = (T)g.ihy, iz = (T)g.ihz;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    T dx[2][2], dy[2][2], dz[2][2];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        dx[b][d] = u[1 + 2 * b + 4 * d][m] - u[2 * b + 4 * d][m];  // edge along x at (ay, az)
        dy[b][d] = u[b + 2 + 4 * d][m] - u[b + 4 * d][m];          // along y at (ax, az)
        dz[b][d] = u[b + 2 * d + 4][m] - u[b + 2 * d][m];          // along z at (ax, ay)
      }
#pragma unroll
    for (int q1 = 0; q1 < 2; ++q1) {
      const T w10 = q1 == 0 ? wp : wm, w11 = q1 == 1 ? wp : wm;
      T tx[2], ty[2], tz[2];
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        tx[d] = w10 * dx[0][d] + w11 * dx[1][d];  // x edges at y point q1, z corner d
        ty[d] = w10 * dy[0][d] + w11 * dy[1][d];  // y edges at x point q1, z corner d
        tz[d] = w10 * dz[0][d] + w11 * dz[1][d];  // z edges at x point q1, y corner d
      }
#pragma unroll
      for (int q2 = 0; q2 < 2; ++q2) {
        const T w20 = q2 == 0 ? wp : wm, w21 = q2 == 1 ? wp : wm;
        gX[q1][q2][m] = ix * (w20 * tx[0] + w21 * tx[1]);
        gY[q1][q2][m] = iy * (w20 * ty[0] + w21 * ty[1]);
        gZ[q1][q2][m] = iz * (w20 * tz[0] + w21 * tz[1]);
      }
    }
  }
}

// Engineering strain at Gauss point (qx, qy, qz) from box_gradients.
template <class T>
__device__ __forceinline__ void box_strain(const T gX[2][2][3], const T gY[2][2][3],
                                           const T gZ[2][2][3], int qx, int qy, int qz, T e[6]) {
  e[0] = gX[qy][qz][0];
  e[1] = gY[qx][qz][1];
  e[2] = gZ[qx][qy][2];
  e[3] = gY[qx][qz][2] + gZ[qx][qy][1];
  e[4] = gX[qy][qz][2] + gZ[qx][qy][0];
  e[5] = gX[qy][qz][1] + gY[qx][qz][0];
}

// Adds the traction rows of the stress s at Gauss point (qx, qy, qz) to the sums
// over the coordinate each gradient factor ignores: SX[qy][qz] over qx of
// (s0, s5, s4), SY[qx][qz] over qy of (s5, s1, s3), SZ[qx][qy] over qz of
// (s4, s3, s2).  FIRST: the point with qx = 0 (SX), qy = 0 (SY), qz = 0 (SZ)
// starts the respective sum.
template <class T>
__device__ __forceinline__ void box_traction_add(const T s[6], int qx, int qy, int qz,
                                                 T SX[2][2][3], T SY[2][2][3], T SZ[2][2][3]) {
  const T rx[3] = {s[0], s[5], s[4]}, ry[3] = {s[5], s[1], s[3]}, rz[3] = {s[4], s[3], s[2]};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    SX[qy][qz][m] = qx == 0 ? rx[m] : SX[qy][qz][m] + rx[m];
    SY[qx][qz][m] = qy == 0 ? ry[m] : SY[qx][qz][m] + ry[m];
    SZ[qx][qy][m] = qz == 0 ? rz[m] : SZ[qx][qy][m] + rz[m];
  }
}

// Element forces f[a][c] = sum_q w detJ (B_q^T s_q)[a][c] from the traction sums
// of box_traction_add: the transposed interpolation back to the corners.
template <class T>
__device__ __forceinline__ void box_forces(const T SX[2][2][3], const T SY[2][2][3],
                                           const T SZ[2][2][3], const BoxGeom& g, T f[8][3]) {
  const T wp = (T)kWp, wm = (T)kWm;
  const T jw = (T)g.jw;
  const T cx = jw * (T)g.ihx, cy = jw * (T)g.ihy, cz = jw * (T)g.ihz;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    T HX[2][2], HY[2][2], HZ[2][2];  // HX[ay][az], HY[ax][az], HZ[ax][ay]
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const T w0 = d == 0 ? wp : wm, w1 = d == 1 ? wp : wm;  // W(point 0|1, corner d)
      T jx[2], jy[2], jz[2];
#pragma unroll
      for (int q1 = 0; q1 < 2; ++q1) {
        jx[q1] = w0 * SX[q1][0][m] + w1 * SX[q1][1][m];
        jy[q1] = w0 * SY[q1][0][m] + w1 * SY[q1][1][m];
        jz[q1] = w0 * SZ[q1][0][m] + w1 * SZ[q1][1][m];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const T v0 = b == 0 ? wp : wm, v1 = b == 1 ? wp : wm;
        HX[b][d] = cx * (v0 * jx[0] + v1 * jx[1]);
        HY[b][d] = cy * (v0 * jy[0] + v1 * jy[1]);
        HZ[b][d] = cz * (v0 * jz[0] + v1 * jz[1]);
      }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
      const T fx = ax ? HX[ay][az] : -HX[ay][az];
      const T fy = ay ? HY[ax][az] : -HY[ax][az];
      const T fz = az ? HZ[ax][ay] : -HZ[ax][ay];
      f[a][m] = fx + fy + fz;
    }
  }
}

// Engineering strain at Gauss point q of a box element from its corner values
// c8[a][off + 0..2] (a = ax + 2 ay + 4 az), factorised: du/dx at q is the 4 edge
// differences along x weighted by W(qy, ay) W(qz, az) (likewise y and z), 84
// operations instead of 8 corners x 9 products plus their factors.  The edge
// differences are formed in the corner values' precision U and rounded to T
// once: with FP64 corners an FP32 strain keeps its relative accuracy when the
// fluctuation varies little across an element (no cancellation in T).
template <class T, class U, int W>
__device__ __forceinline__ void gp_strain_sf(const U (*c8)[W], int off, int q, const BoxGeom& g,
                                             T e[6]) {
  const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
  const T wp = (T)kWp, wm = (T)kWm;
  const T wx[2] = {qx == 0 ? wp : wm, qx == 1 ? wp : wm};  // W(qx, ax = 0 | 1)
  const T wy[2] = {qy == 0 ? wp : wm, qy == 1 ? wp : wm};
  const T wz[2] = {qz == 0 ? wp : wm, qz == 1 ? wp : wm};
  T gx[3] = {(T)0, (T)0, (T)0}, gy[3] = {(T)0, (T)0, (T)0}, gz[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const T wyz = wy[u] * wz[v], wxz = wx[u] * wz[v], wxy = wx[u] * wy[v];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        gx[m] += wyz * (T)(c8[1 + 2 * u + 4 * v][off + m] - c8[2 * u + 4 * v][off + m]);
        gy[m] += wxz * (T)(c8[u + 2 + 4 * v][off + m] - c8[u + 4 * v][off + m]);
        gz[m] += wxy * (T)(c8[u + 2 * v + 4][off + m] - c8[u + 2 * v][off + m]);
      }
    }
  const T ix = (T)g.ihx, iy = (T)g.ihy, iz = (T)g.ihz;
  e[0] = ix * gx[0];
  e[1] = iy * gy[1];
  e[2] = iz * gz[2];
  e[3] = iz * gz[1] + iy * gy[2];
  e[4] = iz * gz[0] + ix * gx[2];
  e[5] = iy * gy[0] + ix * gx[1];
}

// Force at corner a of a box element, w detJ sum_q B_qa^T s_q, from the stresses
// of its 8 Gauss points (s8[q][0..5], q = 4 qx + 2 qy + qz; rows padded to 8).
// Factorised: the x-gradient factor of corner a does not depend on qx, so the x
// rows (s0, s5, s4) of the two points that differ only in qx are added first and
// weighted once (likewise the y rows (s5, s1, s3) over qy and the z rows
// (s4, s3, s2) over qz): ~95 operations instead of the 8-point force rows and an
// 8-lane reduction.
template <class T>
__device__ __forceinline__ void corner_force(const T (*s8)[8], int a, const BoxGeom& g, T f[3]) {
  const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
  const T wp = (T)kWp, wm = (T)kWm;
  const T wx[2] = {ax == 0 ? wp : wm, ax == 1 ? wp : wm};  // W(qx = 0 | 1, ax)
  const T wy[2] = {ay == 0 ? wp : wm, ay == 1 ? wp : wm};
  const T wz[2] = {az == 0 ? wp : wm, az == 1 ? wp : wm};
  T hx[3] = {(T)0, (T)0, (T)0}, hy[3] = {(T)0, (T)0, (T)0}, hz[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      {  // x rows at (qy, qz) = (u, v), summed over qx
        const T* p0 = s8[2 * u + v];
        const T* p1 = s8[4 + 2 * u + v];
        const T w = wy[u] * wz[v];
        hx[0] += w * (p0[0] + p1[0]);
        hx[1] += w * (p0[5] + p1[5]);
        hx[2] += w * (p0[4] + p1[4]);
      }
      {  // y rows at (qx, qz) = (u, v), summed over qy
        const T* p0 = s8[4 * u + v];
        const T* p1 = s8[4 * u + 2 + v];
        const T w = wx[u] * wz[v];
        hy[0] += w * (p0[5] + p1[5]);
        hy[1] += w * (p0[1] + p1[1]);
        hy[2] += w * (p0[3] + p1[3]);
      }
      {  // z rows at (qx, qy) = (u, v), summed over qz
        const T* p0 = s8[4 * u + 2 * v];
        const T* p1 = s8[4 * u + 2 * v + 1];
        const T w = wx[u] * wy[v];
        hz[0] += w * (p0[4] + p1[4]);
        hz[1] += w * (p0[3] + p1[3]);
        hz[2] += w * (p0[2] + p1[2]);
      }
    }
  const T jw = (T)g.jw;
  const T cx = (ax ? jw : -jw) * (T)g.ihx, cy = (ay ? jw : -jw) * (T)g.ihy,
          cz = (az ? jw : -jw) * (T)g.ihz;
#pragma unroll
  for (int m = 0; m < 3; ++m) f[m] = cx * hx[m] + cy * hy[m] + cz * hz[m];
}

// Element (i, j, k) of a periodic nx x ny x nz grid: node classes of its 8 corners.
// (32-bit index arithmetic: an RVE holds fewer than 2^31 elements.)
__device__ __forceinline__ void periodic_corners(int64_t e64, int nx, int ny, int nz, int64_t c[8],
                                                 int& i, int& j, int& k) {
  const unsigned e = (unsigned)e64;
  const unsigned r = e / (unsigned)nx;
  i = (int)(e - r * (unsigned)nx);
  k = (int)(r / (unsigned)ny);
  j = (int)(r - (unsigned)k * (unsigned)ny);
  const int ip = i + 1 == nx ? 0 : i + 1;
  const int jp = j + 1 == ny ? 0 : j + 1;
  const int kp = k + 1 == nz ? 0 : k + 1;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ii = (a & 1) ? ip : i, jj = (a & 2) ? jp : j, kk = (a & 4) ? kp : k;
    c[a] = ii + (int64_t)nx * (jj + (int64_t)ny * kk);
  }
}

// Element that holds periodic node (i, j, k) as its corner a.
// Periodic class of corner a (a = ax + 2 ay + 4 az) of element e.
__device__ __forceinline__ int64_t periodic_corner(int64_t e64, int a, int nx, int ny, int nz) {
  const unsigned e = (unsigned)e64;
  const unsigned r = e / (unsigned)nx;
  int i = (int)(e - r * (unsigned)nx);
  int k = (int)(r / (unsigned)ny);
  int j = (int)(r - (unsigned)k * (unsigned)ny);
  if (a & 1) i = i + 1 == nx ? 0 : i + 1;
  if (a & 2) j = j + 1 == ny ? 0 : j + 1;
  if (a & 4) k = k + 1 == nz ? 0 : k + 1;
  return i + (int64_t)nx * (j + (int64_t)ny * k);
}

__device__ __forceinline__ int64_t periodic_corner_element(int i, int j, int k, int a, int nx,
                                                           int ny, int nz) {
  int ii = i - (a & 1), jj = j - ((a >> 1) & 1), kk = k - ((a >> 2) & 1);
  if (ii < 0) ii += nx;
  if (jj < 0) jj += ny;
  if (kk < 0) kk += nz;
  return ii + (int64_t)nx * (jj + (int64_t)ny * kk);
}

}  // namespace cia
