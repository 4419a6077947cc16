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
#include <math.h>
#include <stdint.h>

#define FV_HD __host__ __device__

/* Quotient and square root of the face physics.  FP64 keeps the correctly rounded
 * operations (the plain form's parity with the reference rests on them).  FP32, the
 * mixed form, takes the hardware reciprocal and square root (about 1 ulp each)
 * instead of the correctly rounded expansions (~10 and ~15 instructions): their error
 * is of the size of the FP32 rounding the mixed form accepts anyway. */
template <typename T> FV_HD __forceinline__ T qdiv(T a, T b) { return a / b; }
template <typename T> FV_HD __forceinline__ T qsqrt(T a) { return sqrt(a); }
template <> FV_HD __forceinline__ float qdiv<float>(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a / b;
#endif
}
template <> FV_HD __forceinline__ float qsqrt<float>(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(a);
#else
  return sqrtf(a);
#endif
}

/* A reciprocal for the reverse pass's derivative factors only (never for a value the
 * forward computes, whose bits the adjoint must reproduce): FP64 from the hardware
 * reciprocal and one Newton step (e = 1 - x r, r + r e: the error squared, within an ulp
 * or two), FP32 the hardware reciprocal itself; the host divides. */
template <typename T> FV_HD __forceinline__ T qrcp(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
  } else {
    return __builtin_amdgcn_rcpf(x);
  }
#else
  return (T)1 / x;
#endif
}

struct DevParams {
  int model, ns, ischeme, limiter, visc_work, neq;
  double gamma, pinf, mu, ymin;
  double baro_nexp, baro_beta, baro_cbar, baro_psat, baro_rsat;
  double mix_RG, mix_Tref, mix_rsat, mix_psat, mix_cliq, mix_cmix;
  double inlet[5];
  double va_eps;
  /* options (fvg_params: iprimtv, wave_speeds, prandtl); read only by the kernels'
   * option variants */
  int iprimtv, pvrs;
  double kheat; /* mu gamma / ((gamma - 1) Pr), 0 without heat conduction */
};

template <int MODEL> struct ModelTraits;
template <> struct ModelTraits<1> { static constexpr int NEQ = 5; };
template <> struct ModelTraits<2> { static constexpr int NEQ = 4; };
template <> struct ModelTraits<3> { static constexpr int NEQ = 5; };

/* primitive face state in compute precision T */
template <typename T> struct FState {
  T r, u[3], e, p, c, y;
  T q[5]; /* conserved vector (SOL::vec) */
  T tv[3]; /* the cell's viscous rows for the face's axis a: tv[i] = vrow(a, i) (vrows) */
  T gt[3]; /* gradient of p / rho (heat conduction option) */
};

/* The viscous flux across an axis-a face uses row a of the stress tensor (flux_viscous.c:
 * 58-73): tau[a][i] = mu (gr[3a+i] + gr[3i+a]) for i != a and mu (2 gr[4a] - 2/3 div gr) for
 * i = a, gr the face mean of the two cells' momentum gradients (gr[3i+j] = d(rho u_i)/dx_j).
 * Both are linear in the gradients, so each cell keeps the six symmetric rows
 *   D_d = 2 g[4d] - (2/3)((g[0] + g[4]) + g[8])   (rows 0, 1, 2)
 *   X_01 = g[1] + g[3], X_02 = g[2] + g[6], X_12 = g[5] + g[7]   (rows 3, 4, 5)
 * and a face forms tau[a][i] = mu ((row of L + row of R) / 2): three values per side
 * instead of seven.  Rows 6..8 hold the gradient of p / rho (heat conduction option).
 * Written without contraction, so that every kernel that forms them rounds the same. */
enum { VROW_HEAT = 6, VROWS = 9 };

FV_HD constexpr int vrow(int a, int i) {
  return a == i ? a : (a + i == 1 ? 3 : (a + i == 2 ? 4 : 5));
}

template <typename T>
FV_HD __forceinline__ void vrows(const T *g, T *out) {
#pragma clang fp contract(off)
  const T div = (g[0] + g[4]) + g[8];
  const T c23 = (T)2.0 / (T)3.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) out[d] = (T)2.0 * g[4 * d] - c23 * div;
  out[3] = g[1] + g[3];
  out[4] = g[2] + g[6];
  out[5] = g[5] + g[7];
}

/* the viscous rows of axis a of a cell from its nine momentum gradients (callers that
 * hold those, as the tangent-linear model and the host checks) */
template <int a, typename T>
FV_HD __forceinline__ void vrows_axis(const T *g, T *tv) {
  T r[6];
  vrows<T>(g, r);
#pragma unroll
  for (int i = 0; i < 3; ++i) tv[i] = r[vrow(a, i)];
}

/* x^e for the liquid branch of the barotropic EOS, where x = rho / rho_sat is a little
 * above 1.  With u = x - 1 (exact for x in [1, 2]) and s = u / (2 + u):
 *   ln x = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + s^6/7 + ...),
 *   x^e  = exp(y) = 1 + y (1 + y/2 (1 + y/3 (1 + ...))),  y = e ln x,
 * both Taylor series (Horner form) of the functions themselves.  For 0 <= y <= 0.1
 * (x <= 1.0164 at e = 6.15, pressures up to about 360 bar above saturation) the atanh
 * series to s^6 leaves a relative error below 1e-17 (s <= 0.0082) and the exp series
 * to y^10 one below 3e-19; the roundings give about 0.6 ulp in all, against the ~200
 * instructions of pow().  Outside that range, or for a non-finite y, it calls pow(). */
FV_HD __forceinline__ double pow_near1(double x, double e) {
  const double u = x - 1.0;
  const double sq = u / (2.0 + u);
  const double s2 = sq * sq;
  double t = fma(s2, 1.0 / 7.0, 1.0 / 5.0);
  t = fma(s2, t, 1.0 / 3.0);
  t = fma(s2, t, 1.0);
  const double y = e * (2.0 * sq * t);
  if (!(y >= 0.0 && y <= 0.1)) return pow(x, e);
  double q = fma(y, 1.0 / 10.0, 1.0);
#pragma unroll
  for (int k = 9; k >= 2; --k) q = fma(y * (1.0 / (double)k), q, 1.0);
  return fma(y, q, 1.0);
}

/* EOS of a conserved vector, always in FP64 (returns false on a failed root) */
template <int MODEL>
FV_HD __forceinline__ bool eos64(const DevParams &P, const double *U, bool clip,
                                      double &r, double u[3], double &e, double &p, double &c,
                                      double &y) {
  /* no fused multiply-add here: the two-phase root -b + sqrt(b^2 - 4d) cancels
   * about three digits, so every product is rounded as in the reference */
#pragma clang fp contract(off)
  /* one reciprocal of rho replaces the reference's four divisions by it */
  r = U[0];
  const double ir = 1.0 / r;
  u[0] = U[1] * ir;
  u[1] = U[2] * ir;
  u[2] = U[3] * ir;
  e = 0.0;
  y = 0.0;
  if constexpr (MODEL == 1) {
    e = U[4];
    double ei = (e * ir) - 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    p = r * (P.gamma - 1.0) * ei - P.gamma * P.pinf;
    c = sqrt(P.gamma * (p + P.pinf) * ir);
    return true;
  } else if constexpr (MODEL == 2) {
    const double rsat = P.baro_rsat;
    if (r >= rsat) {
      /* one power: (r/rsat)^n = (r/rsat)^(n-1) (r/rsat) (the reference calls pow twice;
       * the results agree to an ulp of the power), by series near x = 1 */
      const double x = r / rsat;
      const double xn1 = pow_near1(x, (P.baro_nexp - 1));
      p = P.baro_psat + P.baro_beta * (xn1 * x - 1.0);
      c = sqrt(P.baro_beta * P.baro_nexp * xn1 / rsat);
    } else {
      p = P.baro_psat + P.baro_cbar * ((1.0 / rsat) - (1.0 / r));
      c = sqrt(P.baro_cbar / (r * r));
    }
    return true;
  } else {
    y = U[4] * ir;
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
FV_HD __forceinline__ bool make_state(const DevParams &P, const double *U, bool clip,
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

/* Option bits of the kernels' option variants (fvg_params iprimtv, wave_speeds, prandtl):
 * a variant compiles only the code of its bits, and checks the runtime switches of those
 * bits, so that one instantiation with all bits serves every combination. */
enum { OPT_PRIM = 1, OPT_PVRS = 2, OPT_HEAT = 4, OPT_ALL = 7 };

/* PVRS wave speeds (Toro, Riemann Solvers, 3rd ed., 10.5.2): the star pressure of the
 * linearised Riemann problem, p* = max(0, (pL + pR)/2 - (uR - uL) rho_ avg a_avg / 2)
 * with arithmetic means of density and sound speed, then S_L = uL - aL qL and
 * S_R = uR + aR qR, q_K = 1 where p* <= p_K (rarefaction) and
 * sqrt(1 + (gamma + 1) / (2 gamma) (p* / p_K - 1)) otherwise (the Rankine-Hugoniot shock
 * speed).  Stiffened gas: every pressure is taken as p + p_inf.  uL, uR are the normal
 * velocities. */
template <typename T>
FV_HD __forceinline__ void pvrs_speeds(const DevParams &P, const FState<T> &L, const FState<T> &R,
                                       T uL, T uR, T &SL, T &SR) {
  const T g = (T)P.gamma, pinf = (T)P.pinf;
  const T pL = L.p + pinf, pR = R.p + pinf;
  const T rb = (T)0.5 * (L.r + R.r), ab = (T)0.5 * (L.c + R.c);
  T ps = (T)0.5 * (pL + pR) - (T)0.5 * (uR - uL) * rb * ab;
  ps = ps > (T)0 ? ps : (T)0;
  const T k = qdiv(g + (T)1, (T)2 * g);
  const T qL = ps <= pL ? (T)1 : qsqrt((T)1 + k * (qdiv(ps, pL) - (T)1));
  const T qR = ps <= pR ? (T)1 : qsqrt((T)1 + k * (qdiv(ps, pR) - (T)1));
  SL = uL - L.c * qL;
  SR = uR + R.c * qR;
}

/* Davis-HLLC flux of (L, R) across a face of normal s*e_a, plus the viscous flux;
 * returns out[v] = (F_adv - F_visc).n + HLLC correction (before the -Area factor).
 * OPTS: option bits; OPT_PVRS honours P.pvrs (PVRS wave speeds, model 1), OPT_HEAT
 * P.kheat (heat conduction, model 1 with NS). */
template <int MODEL, bool NS, int a, typename T, int OPTS = 0>
FV_HD __forceinline__ bool hllc_face(const DevParams &P, const FState<T> &L,
                                          const FState<T> &R, T s, T *out) {
  constexpr int NEQ = ModelTraits<MODEL>::NEQ;
  const T zero = (T)0;
  T uL = s * L.u[a], uR = s * R.u[a];
  T ruL = L.r * uL, ruR = R.r * uR;
  T SR = (uL + L.c) > (uR + R.c) ? (uL + L.c) : (uR + R.c);
  T SL = (uL - L.c) < (uR - R.c) ? (uL - L.c) : (uR - R.c);
  if constexpr ((OPTS & OPT_PVRS) && MODEL == 1) {
    if (P.pvrs) pvrs_speeds<T>(P, L, R, uL, uR, SL, SR);
  }
  T SS = qdiv(((L.r * (uL * uL) + L.p) - (R.r * (uR * uR) + R.p) - (SL * ruL) + (SR * ruR)),
              (ruL - ruR - (SL * L.r) + (SR * R.r)));
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
    /* flux_viscous.c:58-73,199-232: face means of velocity and of row a of the stress
     * tensor, from each side's viscous rows of axis a (vrows) */
    T uf[3], ta[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) uf[i] = (T)0.5 * (L.u[i] + R.u[i]);
    const T mu = (T)P.mu + zero;
#pragma unroll
    for (int i = 0; i < 3; ++i) ta[i] = mu * ((T)0.5 * (L.tv[i] + R.tv[i]));
#pragma unroll
    for (int i = 0; i < 3; ++i) F[1 + i] = F[1 + i] - ta[i];
    /* flux_viscous.c:227-232 writes row 4 whenever con->eos == 1, whatever the
     * model: for model 3 that row is the rho Y flux, and the reference subtracts
     * the viscous work from it too */
    if constexpr (MODEL == 1 || MODEL == 3) {
      if (P.visc_work) F[4] = F[4] - (uf[0] * ta[0] + uf[1] * ta[1] + uf[2] * ta[2]);
    }
    /* heat conduction (option): the energy flux loses q_a = -kheat d(p/rho)/dx_a, with the
     * face mean of the two cells' gradients of p/rho */
    if constexpr ((OPTS & OPT_HEAT) && MODEL == 1) {
      if (P.kheat != 0.0) F[4] = F[4] - (T)P.kheat * ((T)0.5 * (L.gt[a] + R.gt[a]));
    }
  }
  T corr[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) corr[v] = zero;
  if (region == 2 || region == 3) {
    /* star_region_euler.c:19-37 */
    T cs = qdiv(Kr * (SK - uK), (SK - SS));
    T Us[NEQ];
    Us[0] = cs;
#pragma unroll
    for (int i = 0; i < 3; ++i) Us[1 + i] = cs * (i == a ? s * SS : Ku[i]);
    if constexpr (MODEL == 1)
      Us[4] = cs * (qdiv(Ke, Kr) + (SS - uK) * (SS + qdiv(Kp, (Kr * (SK - uK)))));
    if constexpr (MODEL == 3) Us[4] = cs * Ky;
#pragma unroll
    for (int v = 0; v < NEQ; ++v) corr[v] = SK * (Us[v] - Kq[v]);
  }
#pragma unroll
  for (int v = 0; v < NEQ; ++v) out[v] = F[v] * s + corr[v];
  return region != 0 && (SS == SS);
}

/* ------------------------------------------------------------------------
 * Derivatives of the face flux (model 1, inviscid) for the discrete adjoint
 * and the tangent-linear model.
 * ---------------------------------------------------------------------- */

/* forward-mode number with one tangent lane (tangent-linear model) */
template <typename T> struct Dual {
  T v, d;
  FV_HD Dual() : v(0), d(0) {}
  FV_HD Dual(double x) : v((T)x), d(0) {}
  FV_HD Dual(T x, T dx) : v(x), d(dx) {}
};
template <typename T> FV_HD inline Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <typename T> FV_HD inline Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <typename T> FV_HD inline Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <typename T> FV_HD inline Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <typename T> FV_HD inline Dual<T> operator/(Dual<T> a, Dual<T> b) {
  T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <typename T> FV_HD inline bool operator<(Dual<T> a, Dual<T> b) { return a.v < b.v; }
template <typename T> FV_HD inline bool operator>(Dual<T> a, Dual<T> b) { return a.v > b.v; }
template <typename T> FV_HD inline bool operator<=(Dual<T> a, Dual<T> b) { return a.v <= b.v; }
template <typename T> FV_HD inline bool operator==(Dual<T> a, Dual<T> b) { return a.v == b.v; }
template <typename T> FV_HD inline Dual<T> sqrt(Dual<T> a) {
  T r = sqrt(a.v);
  return {r, a.d / (T(2) * r)};
}

/* ideal/stiffened-gas face state in any arithmetic T (facercnstr.c:20-35) */
template <typename T>
FV_HD __forceinline__ void make_state_ideal(const DevParams &P, const T *U, FState<T> &s) {
  const T g = (T)P.gamma, pinf = (T)P.pinf;
  s.r = U[0];
  const T ir = qdiv((T)1.0, s.r);
  s.u[0] = U[1] * ir;
  s.u[1] = U[2] * ir;
  s.u[2] = U[3] * ir;
  s.e = U[4];
  T ei = (s.e * ir) - (T)0.5 * (s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]);
  s.p = s.r * (g - (T)1.0) * ei - g * pinf;
  s.c = qsqrt(g * (s.p + pinf) * ir);
  s.y = (T)0.0;
#pragma unroll
  for (int v = 0; v < 5; ++v) s.q[v] = U[v];
}

/* Ideal-gas inversion of the mixed form in FP32.  Returns false where FP32 is
 * not good enough: near vacuum (rho <= 1e-20), where the internal energy is a
 * small difference of large terms (rho e_int < E / 100, or |p| < gamma p_inf /
 * 100), or on a non-finite result.  The caller then rebuilds that face state
 * from the FP64 cell states and inverts it in FP64.  Where FP32 is kept, the
 * relative error of p is about 1e-7 (E / rho e_int) (1 + gamma p_inf / |p|),
 * at most about 1e-5. */
FV_HD __forceinline__ bool make_state_mixed_ideal(const DevParams &P, const float *U,
                                                  FState<float> &s) {
  make_state_ideal<float>(P, U, s);
  /* rho e_int = (p + gamma p_inf) / (gamma - 1) */
  const float gp = (float)(P.gamma * P.pinf), gm1 = (float)(P.gamma - 1.0);
  return s.r > 1e-20f && (s.p + gp) >= 0.01f * gm1 * s.e && fabsf(s.p) >= 0.01f * gp &&
         s.c == s.c && fabsf(s.c) < __builtin_inff();
}

/* Reverse of the viscous part of hllc_face<1, true, a> (flux_viscous.c:58-73,
 * 199-232): F[1+i] -= tau[a][i] and, with visc_work, F[4] -= uf . tau[a], with
 * uf = (uL + uR) / 2 and tau from gr = (gL + gR) / 2 (gradients of the conserved
 * momenta, gr[3 i + j] = d(rho u_i)/dx_j).  Given ob = bar of out (out = F s +
 * corr), adds the velocity part to ULb / URb and returns in gb[9] the bar of
 * each side's gradients (the same for both sides: d gr / d gL = 1/2). */
template <int a, typename T>
FV_HD __forceinline__ void visc_vjp_ideal(const DevParams &P, const T *UL, const T *UR,
                                          const T *gL, const T *gR, T s, const T *ob, T *ULb,
                                          T *URb, T *gb) {
  const T mu = (T)P.mu;
  T gr[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) gr[k] = (T)0.5 * (gL[k] + gR[k]);
  const T div = gr[0] + gr[4] + gr[8];
  T ta[3]; /* tau[a][i] */
#pragma unroll
  for (int i = 0; i < 3; ++i)
    ta[i] = i == a ? mu * ((T)2.0 * gr[4 * a] - ((T)2.0 / (T)3.0) * div)
                   : mu * (gr[3 * a + i] + gr[3 * i + a]);
  const T irL = (T)1 / UL[0], irR = (T)1 / UR[0];
  T uL[3], uR[3], uf[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    uL[i] = UL[1 + i] * irL;
    uR[i] = UR[1 + i] * irR;
    uf[i] = (T)0.5 * (uL[i] + uR[i]);
  }
  T Fb[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) Fb[v] = ob[v] * s;
  const bool work = P.visc_work != 0;
  T tab[3], ufb[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    tab[i] = -Fb[1 + i] - (work ? Fb[4] * uf[i] : (T)0);
    ufb[i] = work ? -Fb[4] * ta[i] : (T)0;
  }
  T grb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) grb[k] = (T)0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i == a) {
      grb[4 * a] += (T)2.0 * mu * tab[i];
      const T d = -((T)2.0 / (T)3.0) * mu * tab[i];
      grb[0] += d;
      grb[4] += d;
      grb[8] += d;
    } else {
      grb[3 * a + i] += mu * tab[i];
      grb[3 * i + a] += mu * tab[i];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) gb[k] = (T)0.5 * grb[k];
  /* uf = (uL + uR) / 2, u_i = q_{1+i} / rho */
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T h = (T)0.5 * ufb[i];
    ULb[1 + i] += h * irL;
    ULb[0] -= h * uL[i] * irL;
    URb[1 + i] += h * irR;
    URb[0] -= h * uR[i] * irR;
  }
}

/* bars of the primitive state (r, u, e, p, c) -> bars of the conserved vector */
template <typename T>
FV_HD __forceinline__ void state_ideal_vjp(const DevParams &P, const FState<T> &s, T rb, T ub[3],
                                           T eb, T pb, T cb, T *Ub) {
  const T g = (T)P.gamma, pinf = (T)P.pinf;
  /* the forward's 1/r (make_state_ideal); the bars multiply by it */
  const T ir = qrcp(s.r);
  /* c = sqrt(q), q = g (p + pinf) ir */
  const T q = g * (s.p + pinf) * ir;
  const T qb = cb * qrcp((T)2 * s.c);
  pb += qb * g * ir;
  rb -= qb * q * ir;
  /* p = r (g - 1) ei - g pinf, ei = e ir - |u|^2 / 2 */
  const T ei = (s.e * ir) - (T)0.5 * (s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]);
  rb += pb * (g - (T)1.0) * ei;
  const T eib = pb * s.r * (g - (T)1.0);
  eb += eib * ir;
  rb -= eib * s.e * ir * ir;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ub[i] -= eib * s.u[i];
    /* u_i = U_{1+i} ir */
    Ub[1 + i] += ub[i] * ir;
    rb -= ub[i] * s.u[i] * ir;
  }
  Ub[0] += rb;
  Ub[4] += eb;
}

/* Hand-written reverse of make_state_ideal + hllc_face<1, false, a> (the same
 * branch choices as the forward): given ob = d(objective)/d(out), accumulate
 * ULb += ob . d out / d UL and URb += ob . d out / d UR. */
template <int a, typename T>
FV_HD __forceinline__ void hllc_vjp_ideal(const DevParams &P, const T *UL, const T *UR, T s,
                                          const T *ob, T *ULb, T *URb) {
  FState<T> L, R;
  make_state_ideal<T>(P, UL, L);
  make_state_ideal<T>(P, UR, R);
  const T zero = (T)0;
  /* forward (hllc_face) */
  const T uL = s * L.u[a], uR = s * R.u[a];
  const T ruL = L.r * uL, ruR = R.r * uR;
  const bool rmaxL = (uL + L.c) > (uR + R.c);
  const T SR = rmaxL ? (uL + L.c) : (uR + R.c);
  const bool lminL = (uL - L.c) < (uR - R.c);
  const T SL = lminL ? (uL - L.c) : (uR - R.c);
  const T num = (L.r * (uL * uL) + L.p) - (R.r * (uR * uR) + R.p) - (SL * ruL) + (SR * ruR);
  const T den = (ruL - ruR - (SL * L.r) + (SR * R.r));
  const T SS = qdiv(num, den);
  int region = 0;
  if (zero <= SL) region = 1;
  else if (SL <= zero && zero <= SS) region = 2;
  else if (SS <= zero && zero <= SR) region = 3;
  else if (SR <= zero) region = 4;
  const bool useL = region <= 2;
  const T Kr = useL ? L.r : R.r, Ke = useL ? L.e : R.e, Kp = useL ? L.p : R.p;
  T Ku[3], Kq[5];
#pragma unroll
  for (int i = 0; i < 3; ++i) Ku[i] = useL ? L.u[i] : R.u[i];
#pragma unroll
  for (int v = 0; v < 5; ++v) Kq[v] = useL ? L.q[v] : R.q[v];
  const T SK = (region == 2) ? SL : SR;
  const T uK = useL ? uL : uR;
  /* reverse */
  T Krb = 0, Keb = 0, Kpb = 0, Kub[3] = {0, 0, 0}, Kqb[5] = {0, 0, 0, 0, 0};
  T uKb = 0, SKb = 0, SSb = 0;
  T Fb[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) Fb[v] = ob[v] * s;
  if (region == 2 || region == 3) {
    /* the forward's star-state values (hllc_face), formed with three reciprocals that
     * the bars multiply by as well: they enter only the derivatives, so their last-bit
     * differences from the forward's quotients do not matter */
    const T W = Kr * (SK - uK);
    const T B = SK - SS;
    const T iKr = qrcp(Kr), iW = qrcp(W), iB = qrcp(B);
    const T cs = W * iB;
    const T Y = SS + Kp * iW;
    const T X = Ke * iKr + (SS - uK) * Y;
    T Us[5];
    Us[0] = cs;
#pragma unroll
    for (int i = 0; i < 3; ++i) Us[1 + i] = cs * (i == a ? s * SS : Ku[i]);
    Us[4] = cs * X;
    T Usb[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      SKb += ob[v] * (Us[v] - Kq[v]);
      Usb[v] = ob[v] * SK;
      Kqb[v] -= ob[v] * SK;
    }
    T csb = Usb[0] + Usb[4] * X;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      csb += Usb[1 + i] * (i == a ? s * SS : Ku[i]);
      if (i == a) SSb += Usb[1 + i] * cs * s;
      else Kub[i] += Usb[1 + i] * cs;
    }
    const T Xb = Usb[4] * cs;
    Keb += Xb * iKr;
    Krb -= Xb * Ke * iKr * iKr;
    SSb += Xb * Y;
    uKb -= Xb * Y;
    const T Yb = Xb * (SS - uK);
    SSb += Yb;
    Kpb += Yb * iW;
    T Wb = -Yb * Kp * iW * iW;
    Wb += csb * iB;
    const T Bb = -csb * cs * iB;
    Krb += Wb * (SK - uK);
    SKb += Wb * Kr;
    uKb -= Wb * Kr;
    SKb += Bb;
    SSb -= Bb;
  }
  /* physical flux of K, component a */
  Kub[a] += Fb[0] * Kr;
  Krb += Fb[0] * Ku[a];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Krb += Fb[1 + i] * Ku[i] * Ku[a];
    Kub[i] += Fb[1 + i] * Kr * Ku[a];
    Kub[a] += Fb[1 + i] * Kr * Ku[i];
  }
  Kpb += Fb[1 + a];
  Kub[a] += Fb[4] * (Ke + Kp);
  Keb += Fb[4] * Ku[a];
  Kpb += Fb[4] * Ku[a];
  /* K -> L or R */
  T Lrb = 0, Leb = 0, Lpb = 0, Lcb = 0, Lub[3] = {0, 0, 0};
  T Rrb = 0, Reb = 0, Rpb = 0, Rcb = 0, Rub[3] = {0, 0, 0};
  T uLb = 0, uRb = 0, SLb = 0, SRb = 0;
  if (useL) {
    Lrb += Krb; Leb += Keb; Lpb += Kpb; uLb += uKb;
#pragma unroll
    for (int i = 0; i < 3; ++i) Lub[i] += Kub[i];
#pragma unroll
    for (int v = 0; v < 5; ++v) ULb[v] += Kqb[v];
  } else {
    Rrb += Krb; Reb += Keb; Rpb += Kpb; uRb += uKb;
#pragma unroll
    for (int i = 0; i < 3; ++i) Rub[i] += Kub[i];
#pragma unroll
    for (int v = 0; v < 5; ++v) URb[v] += Kqb[v];
  }
  if (region == 2) SLb += SKb;
  else if (region == 3) SRb += SKb;
  /* S* = num / den */
  const T iden = qrcp(den);
  const T numb = SSb * iden;
  const T denb = -SSb * SS * iden;
  Lrb += numb * (uL * uL);
  uLb += numb * (T)2 * L.r * uL;
  Lpb += numb;
  Rrb -= numb * (uR * uR);
  uRb -= numb * (T)2 * R.r * uR;
  Rpb -= numb;
  SLb -= numb * ruL;
  T ruLb = -numb * SL;
  SRb += numb * ruR;
  T ruRb = numb * SR;
  ruLb += denb;
  ruRb -= denb;
  SLb -= denb * L.r;
  Lrb -= denb * SL;
  SRb += denb * R.r;
  Rrb += denb * SR;
  Lrb += ruLb * uL;
  uLb += ruLb * L.r;
  Rrb += ruRb * uR;
  uRb += ruRb * R.r;
  if (rmaxL) { uLb += SRb; Lcb += SRb; } else { uRb += SRb; Rcb += SRb; }
  if (lminL) { uLb += SLb; Lcb -= SLb; } else { uRb += SLb; Rcb -= SLb; }
  Lub[a] += s * uLb;
  Rub[a] += s * uRb;
  state_ideal_vjp<T>(P, L, Lrb, Lub, Leb, Lpb, Lcb, ULb);
  state_ideal_vjp<T>(P, R, Rrb, Rub, Reb, Rpb, Rcb, URb);
}

/* limited slope of reconctruct_V3.c with its partials w.r.t. delta0, delta1.  The
 * distances d_cc enter as the reciprocals the forward limiter uses (ip = 1/d_cc(f),
 * im = 1/d_cc(fo); inf on a boundary), so the slopes a, b are the forward's own;
 * the van Albada partials share one reciprocal of the denominator. */
FV_HD __forceinline__ double limited_d(int lim, double eps, bool bnd, double d0, double d1,
                                       double ip, double im, double &g0, double &g1) {
  g0 = 0.0;
  g1 = 0.0;
  if (bnd) return 0.0;
  if (lim == 0) {
    if (fabs(d0) > 1e-10) {
      const double a = -d1 * im, b = d0 * ip;
      if (a * b > 0.0) {
        if (fabs(a) > fabs(b)) { g0 = ip; return b; }
        g1 = -im;
        return a;
      }
    }
    return 0.0;
  }
  const double a = isinf(im) ? 0.0 : -d1 * im;
  const double b = isinf(ip) ? 0.0 : d0 * ip;
  const double N = a * (b * b + eps) + b * (a * a + eps);
  const double D = a * a + b * b + 2.0 * eps;
  const double iD = qrcp(D);
  const double ga = ((b * b + eps) + 2.0 * a * b) * iD - N * 2.0 * a * (iD * iD);
  const double gb = ((a * a + eps) + 2.0 * a * b) * iD - N * 2.0 * b * (iD * iD);
  if (!isinf(ip)) g0 = gb * ip;
  if (!isinf(im)) g1 = -ga * im;
  return N / D;
}

/* The partials of both face states of one axis of a cell: SF[fp] uses
 * g(dP, dM; ip, im) -> (g0, g1), SF[fm] uses g(dM, dP; im, ip) -> (h0, h1).  For van
 * Albada the second slope is the negative of the first (the limiter is symmetric in
 * its two one-sided slopes and odd, and the second call's slopes are the first's,
 * swapped and negated: a' = -b, b' = -a), so h0 = -g1 and h1 = -g0; these are the
 * values the second call computes, bit for bit, from one reciprocal instead of two.
 * Minmod keeps two calls: its 1e-10 gate looks at the first slope only. */
FV_HD __forceinline__ void limited_d_axis(int lim, double eps, bool bp, bool bm, double dP,
                                          double dM, double ip, double im, double &g0,
                                          double &g1, double &h0, double &h1) {
  if (lim == 0) {
    limited_d(lim, eps, bp, dP, dM, ip, im, g0, g1);
    limited_d(lim, eps, bm, dM, dP, im, ip, h0, h1);
    return;
  }
  double c0, c1;
  limited_d(lim, eps, false, dP, dM, ip, im, c0, c1);
  g0 = bp ? 0.0 : c0;
  g1 = bp ? 0.0 : c1;
  h0 = (bm || isinf(im)) ? 0.0 : -c1;
  h1 = (bm || isinf(ip)) ? 0.0 : -c0;
}
