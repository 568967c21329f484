/*
 * fe_stage.hip -- fused RK3 stage kernels of the FastEddy compressible LES for gfx950.
 *
 * One launch evaluates the complete tendency of every prognostic field at
 * every cell and applies the RK3-WS2002 stage update (or, in emit mode, stores
 * the tendency).  A 64x4 work-group owns a tile of columns and marches up the
 * column:
 *   - each thread keeps its own column's 7-deep window (z-3..z+3) of every
 *     field in registers (vertical stencils, z face fluxes computed once and
 *     carried to the next level);
 *   - the tile's xy plane at level z, with a 3-cell halo loaded with periodic
 *     wrap-around, is staged in LDS for the horizontal stencils;
 *   - x and y face quantities (advective fluxes, SGS stresses) are computed
 *     once per face into LDS and differenced by the two cells that share them.
 * Reference path replaced: SRC/HYDRO_CORE/CUDA/cuda_hydroCoreDevice.cu:246-947
 * (tendency build, one launch per phase there) and
 * SRC/TIME_INTEGRATION/CUDA/cuda_RKschemes.cu:21-157 (stage update).
 *
 * Arithmetic is FP32 throughout.  The perturbation pressure uses the
 * well-conditioned form p' = p_BS((1+r)^gamma - 1), r = rho*theta/(rho*theta)_BS - 1
 * (a binomial series for |r| < 1/8) instead of the difference of two ~1e5 Pa
 * powers; amd_pressureForm = 0 selects the reference's difference form.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include "fe_kernels.h"

namespace fe {

constexpr int TX = 64, TY = 4, NT = TX * TY, HH = 3;
constexpr int RW = TX + 2 * HH, RH = TY + 2 * HH;      /* raw plane with 3-cell halo        */
constexpr int DW = TX + 2, DH = TY + 2;                /* derived region [-1,TX] x [-1,TY]  */
constexpr int NHALO = RW * RH - TX * TY;
constexpr int NGX = (TX + 1) * TY, NGY = TX * (TY + 1);

__device__ __forceinline__ int wrapi(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

__device__ __forceinline__ float prof(const StepArgs &a, int p, int z) { return a.prof[p * a.np + z + 3]; }

/* ---- advective face fluxes: u * f(face) from phi[i-3..i+2] (cuda_advectionDevice.cu:123-926) ---- */
__device__ __forceinline__ float weno5_side(float a, float b, float c, float d, float e) {
  const float tol = 1e-6f;
  float fh1 = (1.f / 3.f) * a - (7.f / 6.f) * b + (11.f / 6.f) * c;
  float fh2 = -(1.f / 6.f) * b + (5.f / 6.f) * c + (1.f / 3.f) * d;
  float fh3 = (1.f / 3.f) * c + (5.f / 6.f) * d - (1.f / 6.f) * e;
  float t1 = a - 2.f * b + c, t2 = a - 4.f * b + 3.f * c;
  float t3 = b - 2.f * c + d, t4 = b - d;
  float t5 = c - 2.f * d + e, t6 = 3.f * c - 4.f * d + e;
  float b1 = (13.f / 12.f) * t1 * t1 + 0.25f * t2 * t2;
  float b2 = (13.f / 12.f) * t3 * t3 + 0.25f * t4 * t4;
  float b3 = (13.f / 12.f) * t5 * t5 + 0.25f * t6 * t6;
  float w1 = 0.1f / ((tol + b1) * (tol + b1));
  float w2 = 0.6f / ((tol + b2) * (tol + b2));
  float w3 = 0.3f / ((tol + b3) * (tol + b3));
  float inv = 1.f / (w1 + w2 + w3);
  return (w1 * fh1 + w2 * fh2 + w3 * fh3) * inv;
}
__device__ __forceinline__ float weno3_side(float a, float b, float c) {
  const float tol = 1e-6f;
  float fh1 = -0.5f * a + 1.5f * b, fh2 = 0.5f * b + 0.5f * c;
  float b1 = (b - a) * (b - a), b2 = (c - b) * (c - b);
  float w1 = (1.f / 3.f) / ((tol + b1) * (tol + b1)), w2 = (2.f / 3.f) / ((tol + b2) * (tol + b2));
  return (w1 * fh1 + w2 * fh2) / (w1 + w2);
}

template <int S>
__device__ __forceinline__ float face_flux(float U, float omb, float m3, float m2, float m1, float c0, float p1, float p2) {
  if constexpr (S == 3) {        /* hybrid 5th/6th: b=0 is 5th-order upwind (WS2002) */
    float cen = fmaf(37.f, c0 + m1, fmaf(-8.f, p1 + m2, p2 + m3));
    float upw = fmaf(10.f, c0 - m1, fmaf(-5.f, p1 - m2, p2 - m3));
    float sg = copysignf(omb, U);
    return U * ((cen - sg * upw) * (1.f / 60.f));
  } else if constexpr (S == 2) { /* hybrid 3rd/4th */
    float cen = fmaf(7.f, c0 + m1, -(p1 + m2));
    float upw = fmaf(-3.f, c0 - m1, p1 - m2);
    float sg = copysignf(omb, U);
    return U * ((cen + sg * upw) * (1.f / 12.f));
  } else if constexpr (S == 1) { /* QUICK */
    return 0.125f * (fmaxf(0.f, U) * (6.f * m1 + 3.f * c0 - m2) + fminf(0.f, U) * (6.f * c0 + 3.f * m1 - p1));
  } else if constexpr (S == 4) { /* WENO3 */
    return signbit(U) ? U * weno3_side(p1, c0, m1) : U * weno3_side(m2, m1, c0);
  } else if constexpr (S == 5) { /* WENO5 */
    return signbit(U) ? U * weno5_side(p2, p1, c0, m1, m2) : U * weno5_side(m3, m2, m1, c0, p1);
  } else if constexpr (S == 6) { /* centred 2nd order */
    return U * (0.5f * (c0 + m1));
  } else {                       /* 1st-order upstream */
    return fmaxf(0.f, U) * m1 + fminf(0.f, U) * c0;
  }
}

/* ---- perturbation pressure (cuda_pressureDevice.cu:53-113) ---- */
__device__ __forceinline__ float pprime(const StepArgs &a, float rthm, int z) {
  if (a.pform == 0) return powf(rthm * a.c1p, a.gam) - prof(a, PR_PBSRAW, z);
  float r = (rthm - prof(a, PR_RTH, z)) * prof(a, PR_IRTH, z);
  float s;
  if (fabsf(r) < 0.125f) {
    s = fmaf(r, a.pc[7], a.pc[6]);
    s = fmaf(r, s, a.pc[5]); s = fmaf(r, s, a.pc[4]); s = fmaf(r, s, a.pc[3]);
    s = fmaf(r, s, a.pc[2]); s = fmaf(r, s, a.pc[1]); s = fmaf(r, s, a.pc[0]);
    s = r * s;
  } else {
    s = expm1f(a.gam * log1pf(r));
  }
  return prof(a, PR_PBS, z) * s;
}

/* ---- one column value with the vertical boundary conditions (cuda_BCsDevice.cu:120-241):
 *      rho, rho*theta -> base state; momentum zero-gradient below, geostrophic above (w = 0);
 *      SGS TKE zero-gradient below, 0 above; moisture zero-gradient both ways. ---- */
template <class C>
__device__ __forceinline__ float colval(const StepArgs &a, int f, int z, long col) {
  if (z >= 0 && z < a.Nz) return a.in[f * a.fs + (long)z * a.ps + col];
  if (f == 0) return prof(a, PR_RHO, z);
  if (f == 4) return prof(a, PR_RTH, z);
  if (z < 0) return a.in[f * a.fs + col];
  if (f == 1) return a.uTop;
  if (f == 2) return a.vTop;
  if (f == 3) return 0.f;
  if (C::TKE && f == C::IE) return 0.f;
  return a.in[f * a.fs + (long)(a.Nz - 1) * a.ps + col];
}

template <bool TKE_, int NM_, int SH_, int SE_, int SQV_, int SQL_, bool PHYS_>
struct Cfg {
  static constexpr bool TKE = TKE_;
  static constexpr int NM = NM_;
  static constexpr int NF = 5 + (TKE_ ? 1 : 0) + NM_;
  static constexpr int IE = 5;
  static constexpr int IQV = 5 + (TKE_ ? 1 : 0);
  static constexpr int IQL = IQV + 1;
  static constexpr int SH = SH_, SE = SE_, SQV = SQV_, SQL = SQL_;
  static constexpr bool PHYS = PHYS_;
  static constexpr int NV = 4 + NM_;          /* u, v, w, theta, q/rho ... */
  static constexpr int NB = 4 + NM_;          /* face stress components per face family */
  template <int F> static constexpr int sch() {
    return F < 5 ? SH_ : (TKE_ && F == 5 ? SE_ : (F == IQV ? SQV_ : SQL_));
  }
};

template <class C, int F>
struct FaceLoop {
  __device__ __forceinline__ static void xf(float (*raw)[RH][RW], float (*gx)[TY][TX + 1], int fy, int fx, float U, const float *omb) {
    if constexpr (F < C::NF) {
      const float *r = &raw[F][fy + HH][fx];
      gx[F][fy][fx] = face_flux<C::template sch<F>()>(U, omb[F], r[0], r[1], r[2], r[3], r[4], r[5]);
      FaceLoop<C, F + 1>::xf(raw, gx, fy, fx, U, omb);
    }
  }
  __device__ __forceinline__ static void yf(float (*raw)[RH][RW], float (*gy)[TY + 1][TX], int fy, int fx, float V, const float *omb) {
    if constexpr (F < C::NF) {
      gy[F][fy][fx] = face_flux<C::template sch<F>()>(V, omb[F], raw[F][fy][fx + HH], raw[F][fy + 1][fx + HH], raw[F][fy + 2][fx + HH],
                                                      raw[F][fy + 3][fx + HH], raw[F][fy + 4][fx + HH], raw[F][fy + 5][fx + HH]);
      FaceLoop<C, F + 1>::yf(raw, gy, fy, fx, V, omb);
    }
  }
  __device__ __forceinline__ static void zf(float (*win)[7], float *gz, float W, const float *omb) {
    if constexpr (F < C::NF) {
      gz[F] = face_flux<C::template sch<F>()>(W, omb[F], win[F][1], win[F][2], win[F][3], win[F][4], win[F][5], win[F][6]);
      FaceLoop<C, F + 1>::zf(win, gz, W, omb);
    }
  }
};

/* SGS stress pieces on a face-located cell (cuda_sgsTurbDevice.cu:60-367).  All
 * arguments are velocities (momentum * rho^-1) around the staggered location. */
struct KmCell { float A, B, E; };   /* rho*Km, rho*Km*invPr, (2/3)*rho*e */

/* Surface-layer scheme for one column (cuda_surfaceLayerDevice.cu:112-521). */
__device__ void surface_column(const StepArgs &a, long c2, float rho, float ru, float rv, float rth, float rqv, bool moist, float z1) {
  const SurfArgs &s = a.s;
  float z0 = s.z0m[c2], z0t = s.z0t[c2];
  float z1oz0 = (z1 + z0) / z0, z1ozt0 = (z1 + z0t) / z0t;
  float u1 = ru / rho, v1 = rv / rho;
  float U1 = sqrtf(u1 * u1 + v1 * v1);
  float th1 = rth / rho;
  float q1 = moist ? rqv / rho : 0.f;
  const float kap2 = a.kappa * a.kappa;
  const float pi = 3.14159265358979f;
  float cd = s.cd[c2], ch = s.ch[c2], cq = moist ? s.cq[c2] : 0.f;
  float hf = s.hf[c2], qf = moist ? s.qf[c2] : 0.f, tsk = s.tsk[c2], qsk = moist ? s.qsk[c2] : 0.f;
  float ust = s.ust[c2];
  float th0;
  const float pfac = powf(a.refP / a.presg, a.R_cp);
  if (a.step_it == a.step_itR && a.step_it == 0) {
    float l0 = logf(z1oz0), l1 = logf(z1ozt0);
    cd = kap2 / (l0 * l0);
    ch = kap2 / (l1 * l1);
    cq = kap2 / (l1 * l1);
  } else if (a.step_it == a.step_itR) {
    cd = (ust * ust) / (U1 * U1);
    th0 = tsk * pfac;
    ch = hf / (U1 * (th0 - th1));
    if (moist) cq = qf / (U1 * (qsk - q1));
  }
  if (a.slSel == 1) {
    if (a.slIdeal == 1) {
      float t = a.simTime;
      hf = (t > a.slTs && t <= a.slTe) ? (float)(a.slWth + a.slAmp * sin((double)pi * (t - a.slTs) / (a.slTe - a.slTs))) : a.slWth;
      if (moist) qf = (t > a.slQts && t <= a.slQte) ? (float)(a.slQamp * sin((double)pi * (t - a.slQts) / (a.slQte - a.slQts))) : a.slWq;
    }
  } else if (a.slSel == 2) {
    int tf = (int)roundf(10.0f / a.dt);
    float tff = (float)tf;
    float tskc = tsk, qskc = qsk;
    if ((a.step_it % tf) == 0 && a.stage == 2) {
      float t = a.simTime;
      if (a.slIdeal == 1 && t > a.slTs && t < a.slTe + a.dt * tff) {
        float prev = fmaxf(t - tff * a.dt, a.slTs);
        tskc = tsk + (float)(a.slAmp * (sin((double)pi * (t - a.slTs) / (a.slTe - a.slTs)) - sin((double)pi * (prev - a.slTs) / (a.slTe - a.slTs))));
        if (moist) qskc = qsk + (float)(a.slQamp * (sin((double)pi * (t - a.slQts) / (a.slQte - a.slQts)) - sin((double)pi * (prev - a.slQts) / (a.slQte - a.slQts))));
      } else {
        tskc = (float)(tsk + a.slTr * a.dt * tff / 3600.0);
        if (moist) qskc = (float)(qsk + a.slQr * a.dt * tff / 3600.0);
      }
      tsk = tskc;
      qsk = qskc;
    }
    th0 = tskc * pfac;
    hf = ch * U1 * (th0 - th1);
    if (moist) qf = cq * U1 * (qskc - q1);
  }
  /* Monin-Obukhov iteration: it_max+1 passes */
  th0 = tsk * pfac;
  const float q0 = qsk;
  const int itmax = (a.slStab == 0) ? 5 : 1;
  float cd0 = cd, cdi = cd, chi = ch, cqi = cq, iol = 0.f;
  const float lz = logf(z1oz0), lzt = logf(z1ozt0);
  for (int it = 0; it <= itmax; ++it) {
    float tx = -cd0 * U1 * u1, ty = -cd0 * U1 * v1;
    ust = sqrtf(sqrtf(tx * tx + ty * ty));
    if (a.slStab == 0) {
      if (ust > 0.f) {
        iol = -(a.kappa * a.g * hf) / (ust * ust * ust * th1);
        iol = fmaxf(fminf(iol, 1.f), -1.f);
      } else {
        iol = -1.f;
      }
    } else {
      iol = 0.f;
    }
    float zol = iol * (z1 + z0);
    float pm, ph;
    if (zol < 0.f) {
      float xi = sqrtf(sqrtf(1.f - 16.f * zol));
      float h = 0.5f * (1.f + xi);
      pm = logf(0.5f * (1.f + xi * xi) * h * h) - 2.f * atanf(xi) + 0.5f * pi;
      ph = 2.f * logf(0.5f * (1.f + xi * xi));
    } else {
      pm = -5.f * zol;
      ph = -5.f * zol;
    }
    cdi = kap2 / ((lz - pm) * (lz - pm));
    chi = kap2 / ((lzt - pm) * (lzt - ph));
    cqi = chi;
    if (a.slSel > 1) {
      hf = chi * U1 * (th0 - th1);
      if (moist) qf = cqi * U1 * (q0 - q1);
    }
    cd0 = cdi;
  }
  float Um = sqrtf(u1 * u1 + v1 * v1);
  float t13 = -cdi * Um * ru, t23 = -cdi * Um * rv;
  ust = sqrtf(sqrtf(t13 * t13 + t23 * t23));
  iol = -(a.kappa * a.g * hf) / (ust * ust * ust * th1);
  s.cd[c2] = cdi; s.ch[c2] = chi; s.ust[c2] = ust; s.iol[c2] = iol; s.hf[c2] = hf; s.tsk[c2] = tsk;
  s.t13[c2] = t13; s.t23[c2] = t23; s.th3[c2] = hf * rho;
  if (moist) { s.cq[c2] = cqi; s.qf[c2] = qf; s.qsk[c2] = qsk; s.tq3[c2] = qf * rho; }
}

/* Column pre-pass of the physics stage: SGS length scale (cuda_sgstkeDevice.cu:96-151)
 * and Lilly eddy viscosity Km(z) = c_k/2 (l sqrt(e)|z + l sqrt(e)|z-1)
 * (cuda_sgsTurbDevice.cu:202-254) for z in [-1, Nz], plus the surface layer.
 * Km(-1) = 0 as in the reference, which never computes it. */
template <class C>
__global__ void __launch_bounds__(256) column_kernel(StepArgs a, int do_sgs, int do_surf) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.Nx || y >= a.Ny) return;
  const long col = (long)y * a.P + x;
  if (do_sgs && C::TKE) {
    /* theta and e at z-1, z, z+1 with boundary rules */
    auto th = [&](int z) { return colval<C>(a, 4, z, col) / colval<C>(a, 0, z, col); };
    float thm = th(-2), th0 = th(-1), thp;
    float qprev = 0.f;
    for (int z = -1; z <= a.Nz; ++z) {
      thp = th(z + 1);
      float ri = 1.f / colval<C>(a, 0, z, col);
      float e = colval<C>(a, C::IE, z, col) * ri;
      float dthdz = 0.5f * prof(a, PR_J33, z) * a.dZi * (thp - thm);
      float N2 = (a.g / th0) * dthdz;
      float delta = prof(a, PR_DELTA, z);
      float l = delta;
      float se = sqrtf(e);
      if (N2 > 0.f) l = fmaxf(fminf(0.76f * se * rsqrtf(N2), delta), 1e-2f);
      float q = l * se;
      float km = (z >= 0) ? a.c_k * 0.5f * (q + qprev) : 0.f;
      long o = (long)(z + 1) * a.ps + col;
      a.ls[o] = l;
      a.km[o] = km;
      qprev = q;
      thm = th0;
      th0 = thp;
    }
  }
  if (do_surf) {
    float z1 = 0.5f / (a.dZi * prof(a, PR_J33, 0));
    float rho = a.in[col], ru = a.in[a.fs + col], rv = a.in[2 * a.fs + col], rth = a.in[4 * a.fs + col];
    float rqv = C::NM > 0 ? a.in[C::IQV * a.fs + col] : 0.f;
    surface_column(a, col, rho, ru, rv, rth, rqv, C::NM > 0, z1);
  }
}

/* =============================================================================== */
template <class C>
__global__ void __launch_bounds__(NT) stage_kernel(StepArgs a) {
  constexpr int NF = C::NF;
  __shared__ float raw[NF][RH][RW];
  __shared__ float dvr[DH][DW];
  __shared__ float dvp[DH][DW];
  __shared__ float gx[NF][TY][TX + 1];
  __shared__ float gy[NF][TY + 1][TX];
  /* physics stage only */
  constexpr int NVB = C::PHYS ? 3 : 1;
  constexpr int NVv = C::PHYS ? C::NV : 1;
  __shared__ float vel[NVB][NVv][DH][DW];
  __shared__ float kmp[C::PHYS ? 3 : 1][C::PHYS ? DH : 1][C::PHYS ? DW : 1];
  __shared__ float tb1[C::PHYS ? C::NB : 1][C::PHYS ? TY : 1][C::PHYS ? TX + 1 : 1];
  __shared__ float tb2[C::PHYS ? C::NB : 1][C::PHYS ? TY + 1 : 1][C::PHYS ? TX : 1];

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  const bool active = (x0 + tx < a.Nx) && (y0 + ty < a.Ny);
  const int cx = wrapi(x0 + tx, a.Nx), cy = wrapi(y0 + ty, a.Ny);
  const long col = (long)cy * a.P + cx;
  const int Nz = a.Nz;

  /* horizontal halo cells of this tile: LDS slot and global column, fixed for all levels */
  int hslot[2];
  long hcol[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    int h = tid + p * NT;
    hslot[p] = -1;
    hcol[p] = 0;
    if (h < NHALO) {
      int hx, hy;
      if (h < HH * RW) { hy = h / RW; hx = h % RW; }
      else if (h < 2 * HH * RW) { int hh = h - HH * RW; hy = TY + HH + hh / RW; hx = hh % RW; }
      else { int hh = h - 2 * HH * RW; hy = HH + hh / (2 * HH); int c = hh % (2 * HH); hx = c < HH ? c : TX + c; }
      hslot[p] = hy * RW + hx;
      hcol[p] = (long)wrapi(y0 + hy - HH, a.Ny) * a.P + wrapi(x0 + hx - HH, a.Nx);
    }
  }

  float omb[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    int s = (f < 5) ? C::SH : ((C::TKE && f == C::IE) ? C::SE : (f == C::IQV ? C::SQV : C::SQL));
    float b = (f < 5) ? a.bh : ((C::TKE && f == C::IE) ? a.be : a.bqv);
    omb[f] = (s == 3 || s == 2) ? b : 0.f;
  }

  /* own-column register window: win[f][j] = phi_f(z - 3 + j) */
  float win[NF][7];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int j = 0; j < 7; ++j) win[f][j] = colval<C>(a, f, j - 3, col);

  auto rthm_of = [&](float rth, float rho, float qv) {
    if constexpr (C::NM > 0) return rth * (1.f + a.Rv_Rg * qv / rho * 1e-3f);
    else return rth;
  };
  /* own-column perturbation pressure at z-1, z, z+1 (halo levels take the kMin / kMax-1 value) */
  float pz[3];
  {
    float qv3 = C::NM > 0 ? win[C::IQV][3] : 0.f, qv4 = C::NM > 0 ? win[C::IQV][4] : 0.f;
    pz[1] = pprime(a, rthm_of(win[4][3], win[0][3], qv3), 0);
    pz[0] = pz[1];
    pz[2] = (Nz > 1) ? pprime(a, rthm_of(win[4][4], win[0][4], qv4), 1) : pz[1];
  }
  float gzlo[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) gzlo[f] = 0.f;   /* w_cf = 0 at the ground */
  float rin = 1.f / win[0][3];                 /* rho^-1 at own z */

  /* ---- physics-stage carried state ---- */
  float zc[C::PHYS ? 4 + C::NM : 1];           /* own-column z-face stresses at z: T31, T32, T33, TH3, TQ3... */
  float Aown[3], riz[5];                       /* rho*Km at z-1..z+1; rho^-1 at z-2..z+2 */
  int vb0 = 0, vb1 = 1, vb2 = 2;               /* vel buffers holding z-1, z, z+1 */
  if constexpr (C::PHYS) {
#pragma unroll
    for (int j = 0; j < 5; ++j) riz[j] = 1.f / win[0][j + 1];
    /* vel planes at z=-1 and z=0 for the tile region [-1,TX]x[-1,TY] */
    for (int lev = -1; lev <= 0; ++lev) {
      int b = lev < 0 ? vb0 : vb1;
      for (int i = tid; i < DW * DH; i += NT) {
        int dx = i % DW - 1, dy = i / DW - 1;
        long c = (long)wrapi(y0 + dy, a.Ny) * a.P + wrapi(x0 + dx, a.Nx);
        float r = 1.f / colval<C>(a, 0, lev, c);
        vel[b][0][dy + 1][dx + 1] = colval<C>(a, 1, lev, c) * r;
        vel[b][1][dy + 1][dx + 1] = colval<C>(a, 2, lev, c) * r;
        vel[b][2][dy + 1][dx + 1] = colval<C>(a, 3, lev, c) * r;
        vel[b][3][dy + 1][dx + 1] = colval<C>(a, 4, lev, c) * r;
#pragma unroll
        for (int m = 0; m < C::NM; ++m) vel[b][4 + m][dy + 1][dx + 1] = colval<C>(a, C::IQV + m, lev, c) * r;
      }
    }
    /* Km at z-1 (= 0) and z = 0 for the own column */
    Aown[0] = 0.f;
    Aown[1] = a.km[1 * a.ps + col] * win[0][3];
    __syncthreads();
    /* z-face stresses at z = 0: surface values for T31, T32, TH3 (and TQ3 of qv) */
    {
      const long c2 = col;
      float ls0 = a.ls[1 * a.ps + col];
      float A = Aown[1];
      float Bv = A * (1.f + 2.f * ls0 / prof(a, PR_DELTA, 0));
      float E = C::TKE ? (2.f / 3.f) * win[C::IE][3] : 0.f;
      const int o = 1 + 1 * DW;   /* own cell offset is (ty+1, tx+1) */
      (void)o;
      float w0 = vel[vb1][2][ty + 1][tx + 1], wm = vel[vb0][2][ty + 1][tx + 1];
      float s33 = prof(a, PR_J33, 0) * a.dZi * (w0 - wm);
      float s11 = a.dXi * (vel[vb1][0][ty + 1][tx + 1] - vel[vb1][0][ty + 1][tx]);
      float s22 = a.dYi * (vel[vb1][1][ty + 1][tx + 1] - vel[vb1][1][ty][tx + 1]);
      zc[0] = a.s.t13[c2];
      zc[1] = a.s.t23[c2];
      zc[2] = -2.f * A * s33 + (2.f / 3.f) * A * (s11 + s22 + s33) + E;
      zc[3] = a.s.th3[c2];
#pragma unroll
      for (int m = 0; m < C::NM; ++m) {
        if (m == 0) zc[4] = a.s.tq3[c2];
        else {
          float q0 = vel[vb1][4 + m][ty + 1][tx + 1], qm = vel[vb0][4 + m][ty + 1][tx + 1];
          zc[4 + m] = -Bv * prof(a, PR_J33, 0) * a.dZi * (q0 - qm);
        }
      }
    }
  }

  for (int z = 0; z < Nz; ++z) {
    /* prefetch the window's next level and the RK base */
    float nxt[NF], bse[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) nxt[f] = colval<C>(a, f, z + 4, col);
    if (!a.emitF) {
#pragma unroll
      for (int f = 0; f < NF; ++f) bse[f] = a.base[f * a.fs + (long)z * a.ps + col];
    }
    float kmn = 0.f, lsn = 0.f, lsz = 0.f;
    if constexpr (C::PHYS) {
      kmn = a.km[(long)(z + 2) * a.ps + col];
      lsn = a.ls[(long)(z + 2) * a.ps + col];
      lsz = a.ls[(long)(z + 1) * a.ps + col];
    }

    /* (1) xy plane at level z: own cell from the window, halo ring from memory */
#pragma unroll
    for (int f = 0; f < NF; ++f) raw[f][ty + HH][tx + HH] = win[f][3];
    const long zoff = (long)z * a.ps;
#pragma unroll
    for (int p = 0; p < 2; ++p)
      if (hslot[p] >= 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) (&raw[f][0][0])[hslot[p]] = a.in[f * a.fs + zoff + hcol[p]];
      }
    if constexpr (C::PHYS) {
      /* velocities at z+1 for the tile region (top halo level uses the boundary values) */
      for (int i = tid; i < DW * DH; i += NT) {
        int dx = i % DW - 1, dy = i / DW - 1;
        long c = (long)wrapi(y0 + dy, a.Ny) * a.P + wrapi(x0 + dx, a.Nx);
        float r = 1.f / colval<C>(a, 0, z + 1, c);
        vel[vb2][0][dy + 1][dx + 1] = colval<C>(a, 1, z + 1, c) * r;
        vel[vb2][1][dy + 1][dx + 1] = colval<C>(a, 2, z + 1, c) * r;
        vel[vb2][2][dy + 1][dx + 1] = colval<C>(a, 3, z + 1, c) * r;
        vel[vb2][3][dy + 1][dx + 1] = colval<C>(a, 4, z + 1, c) * r;
#pragma unroll
        for (int m = 0; m < C::NM; ++m) vel[vb2][4 + m][dy + 1][dx + 1] = colval<C>(a, C::IQV + m, z + 1, c) * r;
      }
    }
    __syncthreads();

    /* (2) rho^-1 and p' on [-1,TX]x[-1,TY]; rho*Km, rho*Km*invPr, (2/3)rho*e at level z */
    for (int i = tid; i < DW * DH; i += NT) {
      int dx = i % DW, dy = i / DW;
      float rho = raw[0][dy + HH - 1][dx + HH - 1];
      float qv = C::NM > 0 ? raw[C::IQV][dy + HH - 1][dx + HH - 1] : 0.f;
      dvr[dy][dx] = 1.f / rho;
      dvp[dy][dx] = pprime(a, rthm_of(raw[4][dy + HH - 1][dx + HH - 1], rho, qv), z);
      if constexpr (C::PHYS) {
        long c = (long)wrapi(y0 + dy - 1, a.Ny) * a.P + wrapi(x0 + dx - 1, a.Nx);
        float km = a.km[(long)(z + 1) * a.ps + c], l = a.ls[(long)(z + 1) * a.ps + c];
        float A = rho * km;
        kmp[0][dy][dx] = A;
        kmp[1][dy][dx] = A * (1.f + 2.f * l / prof(a, PR_DELTA, z));
        kmp[2][dy][dx] = C::TKE ? (2.f / 3.f) * raw[C::IE][dy + HH - 1][dx + HH - 1] : 0.f;
      }
    }
    __syncthreads();

    /* (3) x and y face fluxes of every field; SGS stresses on x and y faces */
    {
      const float Dz = prof(a, PR_D, z);
      const float hx = 0.5f * a.dXi * Dz, hy = 0.5f * a.dYi * Dz;
      for (int i = tid; i < NGX + NGY; i += NT) {
        if (i < NGX) {
          int fy = i / (TX + 1), fx = i % (TX + 1);
          float U = hx * (raw[1][fy + HH][fx + HH - 1] * dvr[fy + 1][fx] + raw[1][fy + HH][fx + HH] * dvr[fy + 1][fx + 1]);
          FaceLoop<C, 0>::xf(raw, gx, fy, fx, U, omb);
        } else {
          int j = i - NGX;
          int fy = j / TX, fx = j % TX;
          float V = hy * (raw[2][fy + HH - 1][fx + HH] * dvr[fy][fx + 1] + raw[2][fy + HH][fx + HH] * dvr[fy + 1][fx + 1]);
          FaceLoop<C, 0>::yf(raw, gy, fy, fx, V, omb);
        }
      }
      if constexpr (C::PHYS) {
        const float J33z = prof(a, PR_J33, z);
        const int vz = vb1, vm = vb0;
        for (int i = tid; i < NGX + NGY; i += NT) {
          bool xface = i < NGX;
          int fy, fx;
          if (xface) { fy = i / (TX + 1); fx = i % (TX + 1); }
          else { int j = i - NGX; fy = j / TX; fx = j % TX; }
          /* region coordinates of the cell (fx, fy) and its -x / -y neighbours */
          const int X = fx + 1, Y = fy + 1;
#define VL(b, q, yy, xx) vel[b][q][yy][xx]
          float A = kmp[0][Y][X], Bv = kmp[1][Y][X], E = kmp[2][Y][X];
          float s11 = a.dXi * (VL(vz, 0, Y, X) - VL(vz, 0, Y, X - 1));
          float s22 = a.dYi * (VL(vz, 1, Y, X) - VL(vz, 1, Y - 1, X));
          float s33 = J33z * a.dZi * (VL(vz, 2, Y, X) - VL(vm, 2, Y, X));
          float div = s11 + s22 + s33;
          float s21 = 0.25f * (a.dXi * ((VL(vz, 1, Y, X) - VL(vz, 1, Y, X - 1)) + (VL(vz, 1, Y - 1, X) - VL(vz, 1, Y - 1, X - 1)))
                             + a.dYi * ((VL(vz, 0, Y, X) - VL(vz, 0, Y - 1, X)) + (VL(vz, 0, Y, X - 1) - VL(vz, 0, Y - 1, X - 1))));
          const long c2 = (long)wrapi(y0 + fy, a.Ny) * a.P + wrapi(x0 + fx, a.Nx);
          if (xface) {
            float s31 = 0.25f * (a.dXi * ((VL(vz, 2, Y, X) - VL(vz, 2, Y, X - 1)) + (VL(vm, 2, Y, X) - VL(vm, 2, Y, X - 1)))
                               + J33z * a.dZi * ((VL(vz, 0, Y, X) - VL(vm, 0, Y, X)) + (VL(vz, 0, Y, X - 1) - VL(vm, 0, Y, X - 1))));
            tb1[0][fy][fx] = -2.f * A * s11 + (2.f / 3.f) * A * div + E;
            tb1[1][fy][fx] = -2.f * A * s21;
            tb1[2][fy][fx] = (z == 0) ? a.s.t13[c2] : -2.f * A * s31;
            tb1[3][fy][fx] = -Bv * a.dXi * (VL(vz, 3, Y, X) - VL(vz, 3, Y, X - 1));
#pragma unroll
            for (int m = 0; m < C::NM; ++m) tb1[4 + m][fy][fx] = -Bv * a.dXi * (VL(vz, 4 + m, Y, X) - VL(vz, 4 + m, Y, X - 1));
          } else {
            float s32 = 0.25f * (a.dYi * ((VL(vz, 2, Y, X) - VL(vz, 2, Y - 1, X)) + (VL(vm, 2, Y, X) - VL(vm, 2, Y - 1, X)))
                               + J33z * a.dZi * ((VL(vz, 1, Y, X) - VL(vm, 1, Y, X)) + (VL(vz, 1, Y - 1, X) - VL(vm, 1, Y - 1, X))));
            tb2[0][fy][fx] = -2.f * A * s21;
            tb2[1][fy][fx] = -2.f * A * s22 + (2.f / 3.f) * A * div + E;
            tb2[2][fy][fx] = (z == 0) ? a.s.t23[c2] : -2.f * A * s32;
            tb2[3][fy][fx] = -Bv * a.dYi * (VL(vz, 3, Y, X) - VL(vz, 3, Y - 1, X));
#pragma unroll
            for (int m = 0; m < C::NM; ++m) tb2[4 + m][fy][fx] = -Bv * a.dYi * (VL(vz, 4 + m, Y, X) - VL(vz, 4 + m, Y - 1, X));
          }
#undef VL
        }
      }
    }
    __syncthreads();

    /* (4) own cell: tendencies and the stage update */
    {
      const float Dz = prof(a, PR_D, z), Dp = prof(a, PR_D, z + 1);
      const float J33z = prof(a, PR_J33, z), J33p = prof(a, PR_J33, z + 1);
      const float invD = prof(a, PR_INVD, z);
      /* z face z+1/2 (zero at the lid) */
      float rinp = 1.f / win[0][4];
      float W = (z + 1 >= Nz) ? 0.f : 0.5f * a.dZi * (Dp * J33p * win[3][4] * rinp + Dz * J33z * win[3][3] * rin);
      float gzhi[NF];
      FaceLoop<C, 0>::zf(win, gzhi, W, omb);
      float F[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f)
        F[f] = -invD * ((gx[f][ty][tx + 1] - gx[f][ty][tx]) + (gy[f][ty + 1][tx] - gy[f][ty][tx]) + (gzhi[f] - gzlo[f]));

      const float rho = win[0][3];
      const int X = tx + 1, Y = ty + 1;
      /* pressure gradient (cuda_pressureDevice.cu:118-194) */
      {
        float pzp = pz[2];
        float fac = 1.f;
        if constexpr (C::NM > 0) {
          float qs = 0.f;
#pragma unroll
          for (int m = 0; m < C::NM; ++m) qs += win[C::IQV + m][3];
          fac = 1.f / (1.f + qs * rin * 1e-3f);
        }
        F[1] -= fac * 0.5f * a.dXi * (dvp[Y][X + 1] - dvp[Y][X - 1]);
        F[2] -= fac * 0.5f * a.dYi * (dvp[Y + 1][X] - dvp[Y - 1][X]);
        F[3] -= fac * 0.5f * a.dZi * J33z * (pzp - pz[0]);
      }
      /* Rayleigh damping and buoyancy on rho*w (cuda_rayleighDampingDevice.cu:49-79, cuda_buoyancyDevice.cu:48-67) */
      if (a.damp) F[3] -= prof(a, PR_DAMP, z) * win[3][3];
      if constexpr (C::NM == 0) {
        F[3] -= a.g * (rho - prof(a, PR_RHO, z));
      } else {
        float rt = rho;
#pragma unroll
        for (int m = 0; m < C::NM; ++m) rt += win[C::IQV + m][3] * 1e-3f;
        F[3] -= (rho / rt) * a.g * (rt - prof(a, PR_RHO, z));
      }
      /* Coriolis with geostrophic forcing (cuda_coriolisDevice.cu:50-58) */
      {
        float u = win[1][3] * rin, v = win[2][3] * rin, w = win[3][3] * rin;
        F[1] += a.fH * (v - prof(a, PR_VG, z)) - a.fV * w;
        F[2] -= a.fH * (u - prof(a, PR_UG, z));
        F[3] += a.fV * (u - prof(a, PR_UG, z));
      }
      /* 6th-order vertical filter on rho*w (cuda_filtersDevice.cu:75-122) */
      if (a.filt != 0.f) {
        const float *w = win[3];
        float gm = w[3] - w[2], gp = w[4] - w[3];
        float fm = 10.f * (w[3] - w[2]) - 5.f * (w[4] - w[1]) + (w[5] - w[0]);
        float fp = 10.f * (w[4] - w[3]) - 5.f * (w[5] - w[2]) + (w[6] - w[1]);
        fm = signbit(gm * fm) ? 0.f : fm;
        fp = signbit(gp * fp) ? 0.f : fp;
        F[3] += a.filt * (fp - fm);
      }

      if constexpr (C::PHYS) {
        /* z-face stresses at z+1 for the own column */
        float zn[4 + C::NM];
        {
          const int vz = vb1, vp = vb2;
          float A = kmn * win[0][4];
          float lsd = prof(a, PR_DELTA, z + 1);
          float Bv = A * (1.f + 2.f * lsn / lsd);
          float E = C::TKE ? (2.f / 3.f) * win[C::IE][4] : 0.f;
          float wpY = vel[vp][2][Y][X], wzY = vel[vz][2][Y][X];
          float s33 = J33p * a.dZi * (wpY - wzY);
          float s11 = a.dXi * (vel[vp][0][Y][X] - vel[vp][0][Y][X - 1]);
          float s22 = a.dYi * (vel[vp][1][Y][X] - vel[vp][1][Y - 1][X]);
          float s31 = 0.25f * (a.dXi * ((vel[vp][2][Y][X] - vel[vp][2][Y][X - 1]) + (vel[vz][2][Y][X] - vel[vz][2][Y][X - 1]))
                             + J33p * a.dZi * ((vel[vp][0][Y][X] - vel[vz][0][Y][X]) + (vel[vp][0][Y][X - 1] - vel[vz][0][Y][X - 1])));
          float s32 = 0.25f * (a.dYi * ((vel[vp][2][Y][X] - vel[vp][2][Y - 1][X]) + (vel[vz][2][Y][X] - vel[vz][2][Y - 1][X]))
                             + J33p * a.dZi * ((vel[vp][1][Y][X] - vel[vz][1][Y][X]) + (vel[vp][1][Y - 1][X] - vel[vz][1][Y - 1][X])));
          zn[0] = -2.f * A * s31;
          zn[1] = -2.f * A * s32;
          zn[2] = -2.f * A * s33 + (2.f / 3.f) * A * (s11 + s22 + s33) + E;
          zn[3] = -Bv * J33p * a.dZi * (vel[vp][3][Y][X] - vel[vz][3][Y][X]);
#pragma unroll
          for (int m = 0; m < C::NM; ++m) zn[4 + m] = -Bv * J33p * a.dZi * (vel[vp][4 + m][Y][X] - vel[vz][4 + m][Y][X]);
          Aown[2] = A;
        }
        /* divergence of the SGS stresses (cuda_sgsTurbDevice.cu:447-547) */
        const float zdz = J33z * a.dZi;
        F[1] -= a.dXi * (tb1[0][ty][tx + 1] - tb1[0][ty][tx]) + a.dYi * (tb2[0][ty + 1][tx] - tb2[0][ty][tx]) + zdz * (zn[0] - zc[0]);
        F[2] -= a.dXi * (tb1[1][ty][tx + 1] - tb1[1][ty][tx]) + a.dYi * (tb2[1][ty + 1][tx] - tb2[1][ty][tx]) + zdz * (zn[1] - zc[1]);
        F[3] -= a.dXi * (tb1[2][ty][tx + 1] - tb1[2][ty][tx]) + a.dYi * (tb2[2][ty + 1][tx] - tb2[2][ty][tx]) + zdz * (zn[2] - zc[2]);
        F[4] -= a.dXi * (tb1[3][ty][tx + 1] - tb1[3][ty][tx]) + a.dYi * (tb2[3][ty + 1][tx] - tb2[3][ty][tx]) + zdz * (zn[3] - zc[3]);
#pragma unroll
        for (int m = 0; m < C::NM; ++m)
          F[C::IQV + m] -= a.dXi * (tb1[4 + m][ty][tx + 1] - tb1[4 + m][ty][tx]) + a.dYi * (tb2[4 + m][ty + 1][tx] - tb2[4 + m][ty][tx])
                          + zdz * (zn[4 + m] - zc[4 + m]);
        if constexpr (C::TKE) {
          /* prognostic SGS TKE sources (cuda_sgstkeDevice.cu:155-397) */
          const float rhoe = win[C::IE][3];
          const float e = rhoe * rin;
          const float th = win[4][3] * rin;
          float src = a.g / th * zc[3];
          {
            float ce = 1.9f * a.c_k + (0.93f - 1.9f * a.c_k) * lsz / prof(a, PR_DELTA, z);
            src += -ce * rhoe * sqrtf(e) / lsz;
          }
          {
            const int vz = vb1;
            float gxu = 0.5f * a.dXi * (vel[vz][0][Y][X + 1] - vel[vz][0][Y][X - 1]);
            float gyu = 0.5f * a.dYi * (vel[vz][0][Y + 1][X] - vel[vz][0][Y - 1][X]);
            float gzu = 0.5f * zdz * (vel[vb2][0][Y][X] - vel[vb0][0][Y][X]);
            float gxv = 0.5f * a.dXi * (vel[vz][1][Y][X + 1] - vel[vz][1][Y][X - 1]);
            float gyv = 0.5f * a.dYi * (vel[vz][1][Y + 1][X] - vel[vz][1][Y - 1][X]);
            float gzv = 0.5f * zdz * (vel[vb2][1][Y][X] - vel[vb0][1][Y][X]);
            float gxw = 0.5f * a.dXi * (vel[vz][2][Y][X + 1] - vel[vz][2][Y][X - 1]);
            float gyw = 0.5f * a.dYi * (vel[vz][2][Y + 1][X] - vel[vz][2][Y - 1][X]);
            float gzw = 0.5f * zdz * (vel[vb2][2][Y][X] - vel[vb0][2][Y][X]);
            float t11 = tb1[0][ty][tx], t21 = tb1[1][ty][tx], t31 = tb1[2][ty][tx];
            float t22 = tb2[1][ty][tx], t32 = tb2[2][ty][tx], t33 = zc[2];
            src -= t11 * gxu + t21 * gyu + t31 * gzu + t21 * gxv + t22 * gyv + t32 * gzv + t31 * gxw + t32 * gyw + t33 * gzw;
          }
          {
            /* turbulent transport: d/dx_j (rho Km de/dx_j), centred gradients */
            const int R = ty + HH, Cc = tx + HH;
            auto ecell = [&](int yy, int xx) { return raw[C::IE][yy][xx] / raw[0][yy][xx]; };
            float ex2p = ecell(R, Cc + 2), ex2m = ecell(R, Cc - 2), ey2p = ecell(R + 2, Cc), ey2m = ecell(R - 2, Cc);
            float dexp = 0.5f * a.dXi * (ex2p - e), dexm = 0.5f * a.dXi * (e - ex2m);
            float deyp = 0.5f * a.dYi * (ey2p - e), deym = 0.5f * a.dYi * (e - ey2m);
            float ezp2 = win[C::IE][5] * riz[4], ezm2 = win[C::IE][1] * riz[0];
            float dezp = 0.5f * J33p * a.dZi * (ezp2 - e);
            float dezm = 0.5f * prof(a, PR_J33, z - 1) * a.dZi * (e - ezm2);
            src += a.dXi * (dexp * kmp[0][Y][X + 1] - dexm * kmp[0][Y][X - 1])
                 + a.dYi * (deyp * kmp[0][Y + 1][X] - deym * kmp[0][Y - 1][X])
                 + zdz * (dezp * Aown[2] - dezm * Aown[0]);
          }
          F[C::IE] += src;
        }
        if constexpr (C::NM > 1) {
          /* saturation adjustment and latent heating (cuda_moistureDevice.cu:98-245) */
          if (a.mpActive) {
            const float rth = win[4][3];
            const float thd = rth * rin;
            const float c1 = a.c1p;
            float pr = a.condBase ? powf(prof(a, PR_RTH, z) * c1, a.gam) : powf(rth * c1, a.gam);
            float Td = thd * powf(pr / a.refP, a.R_cp);
            float Tc = fmaxf(-80.f, Td - 273.16f);
            float pvs;
            if (Td >= 273.15f)
              pvs = (6.11239921f + Tc * (0.443987641f + Tc * (0.142986287e-1f + Tc * (0.264847430e-3f + Tc * (0.302950461e-5f
                    + Tc * (0.206739458e-7f + Tc * (0.640689451e-10f + Tc * (-0.952447341e-13f + -0.976195544e-15f * Tc)))))))) * 100.f;
            else
              pvs = (6.11147274f + Tc * (0.503160820f + Tc * (0.188439774e-1f + Tc * (0.420895665e-3f + Tc * (0.615021634e-5f
                    + Tc * (0.602588177e-7f + Tc * (0.385852041e-9f + Tc * (0.146898966e-11f + 0.252751365e-14f * Tc)))))))) * 100.f;
            pvs = fminf(0.99f * pr, pvs);
            const float qv = win[C::IQV][3], ql = win[C::IQL][3];
            float ffor = 0.f, flim = ql * 1e-3f;
            if (a.condSel == 1) {
              ffor = qv * 1e-3f - pvs / (a.R_vapor * Td);
            } else {
              float rvs = (a.condSel == 2) ? pvs / (a.R_vapor * Td) : (pvs / (pr - pvs) / a.Rv_Rg) * rho;
              ffor = (qv * 1e-3f - rvs) / (1.f + (a.L_v * a.L_v * rvs * rin / (a.cp * a.R_vapor * Td * Td)));
            }
            float fc = (fmaxf(ffor, 0.f) - fmaxf(fminf(flim, -ffor), 0.f)) / a.condTs;
            F[C::IQV] -= fc * 1e3f;
            F[C::IQL] += fc * 1e3f;
            float prf = powf(rth * c1, a.gam);
            float Tf = thd * powf(prf / a.refP, a.R_cp);
            F[4] += ((thd * a.L_v) / (Tf * a.cp)) * fc;
          }
        }
        /* carry z-face stresses and rho*Km up one level */
#pragma unroll
        for (int q = 0; q < 4 + C::NM; ++q) zc[q] = zn[q];
        Aown[0] = Aown[1];
        Aown[1] = Aown[2];
      }

      if (active) {
        const long o = (long)z * a.ps + col;
        if (a.emitF) {
#pragma unroll
          for (int f = 0; f < NF; ++f) a.out[f * a.fs + o] = F[f];
        } else {
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            float v = fmaf(a.coefdt, F[f], bse[f]);
            if (C::TKE && f == C::IE) v = fmaxf(v, 1e-10f);
            if (C::NM > 0 && f >= C::IQV) v = fmaxf(v, 0.f);
            a.out[f * a.fs + o] = v;
          }
        }
        if (a.pout) a.pout[o] = pz[1];
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) gzlo[f] = gzhi[f];
      rin = rinp;
    }

    /* (5) advance the column window and the own-column pressure */
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
      for (int j = 0; j < 6; ++j) win[f][j] = win[f][j + 1];
      win[f][6] = nxt[f];
    }
    pz[0] = pz[1];
    pz[1] = pz[2];
    if (z + 2 < Nz) {
      float qv = C::NM > 0 ? win[C::IQV][4] : 0.f;
      pz[2] = pprime(a, rthm_of(win[4][4], win[0][4], qv), z + 2);
    }
    if constexpr (C::PHYS) {
#pragma unroll
      for (int j = 0; j < 4; ++j) riz[j] = riz[j + 1];
      riz[4] = 1.f / win[0][5];
      int t = vb0; vb0 = vb1; vb1 = vb2; vb2 = t;
    }
    __syncthreads();
  }
}

}  // namespace fe
