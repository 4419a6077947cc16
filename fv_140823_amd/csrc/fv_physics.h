/*
 * Device physics of the face flux: equation of state, HLLC with Davis wave
 * speeds and the viscous stress, written once for FP64 (plain form) and FP32
 * (mixed form).  The formulas and their operation order follow the reference:
 *   ideal / stiffened gas   facercnstr.c:30-35
 *   barotropic (model 2)    barotropic_eos.c:3-30
 *   two-phase (model 3)     barotropic_eos.c:32-129, clip facercnstr.c:43-44
 *   wave speeds, S*         compute_wave_speeds.c:33-82
 *   HLLC regions            hllc_euler.c:18-50
 *   star state              star_region_euler.c:19-37
 *   physical flux           flux_adv.c:13-43
 *   viscous flux            flux_viscous.c:58-73,199-232 (turb = 0)
 * On the octree every face normal is +-e_a and the tangents are +-e_b, so the
 * reference's rotations by (n, t1, t2) (compute_wave_speeds.c:33-43,
 * rotate_vectors.c:32-91, star_region_euler.c:27-30) reduce exactly to picking
 * component a with sign s; the products with the zero entries only add signed
 * zeros.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct DevParams {
  int model, ns, ischeme, limiter, visc_work, neq;
  double gamma, pinf, mu, ymin;
  double baro_nexp, baro_beta, baro_cbar, baro_psat, baro_rsat;
  double mix_RG, mix_Tref, mix_rsat, mix_psat, mix_cliq, mix_cmix;
  double inlet[5];
  double va_eps;
};

template <int MODEL> struct ModelTraits;
template <> struct ModelTraits<1> { static constexpr int NEQ = 5; };
template <> struct ModelTraits<2> { static constexpr int NEQ = 4; };
template <> struct ModelTraits<3> { static constexpr int NEQ = 5; };

/* primitive face state in compute precision T */
template <typename T> struct FState {
  T r, u[3], e, p, c, y;
  T q[5]; /* conserved vector (SOL::vec) */
  T g[9]; /* gradients of rho u, rho v, rho w (SG rows 1..3) */
};

/* EOS of a conserved vector, always in FP64 (returns false on a failed root) */
template <int MODEL>
__device__ __forceinline__ bool eos64(const DevParams &P, const double *U, bool clip,
                                      double &r, double u[3], double &e, double &p, double &c,
                                      double &y) {
  r = U[0];
  u[0] = U[1] / r;
  u[1] = U[2] / r;
  u[2] = U[3] / r;
  e = 0.0;
  y = 0.0;
  if constexpr (MODEL == 1) {
    e = U[4];
    double ei = (e / r) - 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    p = r * (P.gamma - 1.0) * ei - P.gamma * P.pinf;
    c = sqrt(P.gamma * (p + P.pinf) / r);
    return true;
  } else if constexpr (MODEL == 2) {
    const double rsat = P.baro_rsat;
    if (r >= rsat) {
      p = P.baro_psat + P.baro_beta * (pow((r / rsat), P.baro_nexp) - 1.0);
      c = sqrt(P.baro_beta * P.baro_nexp * (pow((r / rsat), (P.baro_nexp - 1))) / rsat);
    } else {
      p = P.baro_psat + P.baro_cbar * ((1.0 / rsat) - (1.0 / r));
      c = sqrt(P.baro_cbar / (r * r));
    }
    return true;
  } else {
    y = U[4] / r;
    if (clip) {
      if (y < (P.ymin / 1000.0)) y = (P.ymin / 1000.0);
      if (y > (1.0 - (P.ymin / 1000.0))) y = 1.0 - (P.ymin / 1000.0);
    }
    double ct = P.mix_cliq;
    double b = ct * ct * P.mix_rsat - P.mix_psat - r * y * P.mix_RG * P.mix_Tref -
               r * (ct * ct) * (1.0 - y);
    double d = (P.mix_psat - ct * ct * P.mix_rsat) * r * y * P.mix_RG * P.mix_Tref;
    double disc = b * b - 4.0 * d;
    if (!(sqrt(disc) == sqrt(disc))) {
      ct = P.mix_cmix;
      b = ct * ct * P.mix_rsat - P.mix_psat - r * y * P.mix_RG * P.mix_Tref -
          r * (ct * ct) * (1.0 - y);
      d = (P.mix_psat - ct * ct * P.mix_rsat) * r * y * P.mix_RG * P.mix_Tref;
      disc = b * b - 4.0 * d;
      if (!(sqrt(disc) == sqrt(disc))) { p = c = __builtin_nan(""); return false; }
    }
    double p1 = (-b + sqrt(disc)) / 2.0;
    double p2 = (-b - sqrt(disc)) / 2.0;
    p = p1 > p2 ? p1 : p2;
    c = ct;
    return true;
  }
}

/* face state: EOS in FP64, then the primitive state in compute precision T */
template <int MODEL, typename T>
__device__ __forceinline__ bool make_state(const DevParams &P, const double *U, bool clip,
                                           FState<T> &s) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  double r, u[3], e, p, c, y;
  bool ok = eos64<MODEL>(P, U, clip, r, u, e, p, c, y);
  s.r = (T)r;
  s.u[0] = (T)u[0];
  s.u[1] = (T)u[1];
  s.u[2] = (T)u[2];
  s.e = (T)e;
  s.p = (T)p;
  s.c = (T)c;
  s.y = (T)y;
#pragma unroll
  for (int v = 0; v < NEQ; ++v) s.q[v] = (T)U[v];
  return ok;
}

/* Davis-HLLC flux of (L, R) across a face of normal s*e_a, plus the viscous flux;
 * returns out[v] = (F_adv - F_visc).n + HLLC correction (before the -Area factor). */
template <int MODEL, bool NS, int a, typename T>
__device__ __forceinline__ bool hllc_face(const DevParams &P, const FState<T> &L,
                                          const FState<T> &R, T s, T *out) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const T zero = (T)0;
  T uL = s * L.u[a], uR = s * R.u[a];
  T ruL = L.r * uL, ruR = R.r * uR;
  T SR = (uL + L.c) > (uR + R.c) ? (uL + L.c) : (uR + R.c);
  T SL = (uL - L.c) < (uR - R.c) ? (uL - L.c) : (uR - R.c);
  T SS = ((L.r * (uL * uL) + L.p) - (R.r * (uR * uR) + R.p) - (SL * ruL) + (SR * ruR)) /
         (ruL - ruR - (SL * L.r) + (SR * R.r));
  int region = 0;
  if (zero <= SL) region = 1;
  else if (SL <= zero && zero <= SS) region = 2;
  else if (SS <= zero && zero <= SR) region = 3;
  else if (SR <= zero) region = 4;
  /* state K of the flux (L for regions 1-2, R for 3-4), picked field by field so
   * both states stay in registers */
  const bool useL = region <= 2;
  const T Kr = useL ? L.r : R.r, Ke = useL ? L.e : R.e, Kp = useL ? L.p : R.p;
  const T Ky = useL ? L.y : R.y;
  T Ku[3], Kq[NEQ];
#pragma unroll
  for (int i = 0; i < 3; ++i) Ku[i] = useL ? L.u[i] : R.u[i];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) Kq[v] = useL ? L.q[v] : R.q[v];
  T SK = (region == 2) ? SL : SR;
  T uK = useL ? uL : uR;
  /* physical flux of state K, normal component a (flux_adv.c:13-43) */
  T F[NEQ];
  F[0] = Ku[a] * Kr;
#pragma unroll
  for (int i = 0; i < 3; ++i) F[1 + i] = Kr * Ku[i] * Ku[a] + (i == a ? Kp : zero);
  if constexpr (MODEL == 1) F[4] = Ku[a] * (Ke + Kp);
  if constexpr (MODEL == 3) F[4] = Ku[a] * Kr * Ky;
  if constexpr (NS) {
    /* flux_viscous.c:58-73,199-232: face means of velocity and of the gradients
     * of the conserved momenta (SG rows 1..3) */
    T uf[3], gr[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) uf[i] = (T)0.5 * (L.u[i] + R.u[i]);
#pragma unroll
    for (int k = 0; k < 9; ++k) gr[k] = (T)0.5 * (L.g[k] + R.g[k]);
    const T mu = (T)P.mu + zero;
    /* gr[3*i + j] = d(rho u_i)/dx_j */
    T div = gr[0] + gr[4] + gr[8];
    T tau[3][3];
    tau[0][0] = mu * ((T)2.0 * gr[0] - ((T)2.0 / (T)3.0) * div);
    tau[1][1] = mu * ((T)2.0 * gr[4] - ((T)2.0 / (T)3.0) * div);
    tau[2][2] = mu * ((T)2.0 * gr[8] - ((T)2.0 / (T)3.0) * div);
    tau[0][1] = tau[1][0] = mu * (gr[1] + gr[3]);
    tau[0][2] = tau[2][0] = mu * (gr[2] + gr[6]);
    tau[1][2] = tau[2][1] = mu * (gr[5] + gr[7]);
#pragma unroll
    for (int i = 0; i < 3; ++i) F[1 + i] = F[1 + i] - tau[a][i];
    if constexpr (MODEL == 1) {
      if (P.visc_work) F[4] = F[4] - (uf[0] * tau[a][0] + uf[1] * tau[a][1] + uf[2] * tau[a][2]);
    }
  }
  T corr[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) corr[v] = zero;
  if (region == 2 || region == 3) {
    /* star_region_euler.c:19-37 */
    T cs = Kr * (SK - uK) / (SK - SS);
    T Us[NEQ];
    Us[0] = cs;
#pragma unroll
    for (int i = 0; i < 3; ++i) Us[1 + i] = cs * (i == a ? s * SS : Ku[i]);
    if constexpr (MODEL == 1)
      Us[4] = cs * ((Ke / Kr) + (SS - uK) * (SS + (Kp / (Kr * (SK - uK)))));
    if constexpr (MODEL == 3) Us[4] = cs * Ky;
#pragma unroll
    for (int v = 0; v < NEQ; ++v) corr[v] = SK * (Us[v] - Kq[v]);
  }
#pragma unroll
  for (int v = 0; v < NEQ; ++v) out[v] = F[v] * s + corr[v];
  return region != 0 && (SS == SS);
}
