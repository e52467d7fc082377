// Per-point pieces of the Π pipeline shared by the engines that filter planes: the velocity in the
// frame the filter works in, the plane set (components, then their products) and the epilogue that
// turns filtered numerators into ū and τ (Diagnostics.jl:496-650, 825-846). Device code only.
#pragma once
#include <hip/hip_runtime.h>

#include "cgef_launch.hpp"

namespace cgef {

// Velocity in the filter's frame: planetary Cartesian on the sphere (Diagnostics.jl:555-573,
// e_east = (−sinλ, cosλ, 0), e_north = (−sinφcosλ, −sinφsinλ, cosφ)), as given otherwise.
__device__ __forceinline__ void frame_vel(const GeomArgs &G, int dtype, const void *u, const void *v, const void *w,
                                          int64_t off, int64_t t, int64_t i, int64_t j, double U[3]) {
    const double uu = dtype == CGEF_F32 ? (double)((const float *)u)[off + t] : ((const double *)u)[off + t];
    const double vv = dtype == CGEF_F32 ? (double)((const float *)v)[off + t] : ((const double *)v)[off + t];
    if (G.mode == 1) {
        const double sl = G.trig[i], cl = G.trig[G.n[0] + i];
        const double sp = G.trig[2 * G.n[0] + j], cp = G.trig[2 * G.n[0] + G.n[1] + j];
        U[0] = -sl * uu - sp * cl * vv;
        U[1] = cl * uu - sp * sl * vv;
        U[2] = cp * vv;
    } else {
        U[0] = uu;
        U[1] = vv;
        U[2] = G.mode == 2 ? (dtype == CGEF_F32 ? (double)((const float *)w)[off + t] : ((const double *)w)[off + t]) : 0.0;
    }
}

// Plane f of the Π set at a point whose (masked) frame velocity is q: 2-D Cartesian q0 q1 | q0q0 q0q1
// q1q1; otherwise q0 q1 q2 | q0q0 q0q1 q0q2 q1q1 q1q2 q2q2.
__device__ __forceinline__ double pi_plane_value(int mode, int f, const double q[3]) {
    const int NC = mode == 0 ? 2 : 3;
    if (f < NC) return q[f];
    f -= NC;
    if (NC == 2) return f == 0 ? q[0] * q[0] : (f == 1 ? q[0] * q[1] : q[1] * q[1]);
    const int a = f < 3 ? 0 : (f < 5 ? 1 : 2);
    const int b = f < 3 ? f : (f < 5 ? f - 2 : 2);
    return q[a] * q[b];
}

// ū = F[q]·inv, τ = F[q_a q_b]·inv − ū_a ū_b where `ok` (0 elsewhere), rotated to the local (east,
// north) frame on the sphere; writes ū (FP64 planes the strain reads), τ (3 or 6 planes of T) and,
// when asked for, ū/v̄/w̄ in T. F holds the plane numerators in pi_plane_value order; inv is the
// reciprocal of the normalization.
template <typename T>
__device__ __forceinline__ void pi_point_epilogue_inv(const GeomArgs &G, int64_t t, int64_t i, int64_t j,
                                                      const double *F, bool ok, double inv, int64_t b,
                                                      double *__restrict__ ubar, T *__restrict__ tau, T *__restrict__ uo,
                                                      T *__restrict__ vo, T *__restrict__ wo) {
    const int64_t P = G.npts, Nx = G.n[0], Ny = G.n[1];
    const int NC = G.mode == 0 ? 2 : 3;
    double U[3] = {0.0, 0.0, 0.0}, tc[6] = {0, 0, 0, 0, 0, 0};
    if (ok) {
        // every index below is a compile-time constant, so F and the temporaries stay in registers
        if (NC == 2) {
            U[0] = F[0] * inv;
            U[1] = F[1] * inv;
            tc[0] = F[2] * inv - U[0] * U[0];
            tc[1] = F[3] * inv - U[0] * U[1];
            tc[3] = F[4] * inv - U[1] * U[1];
        } else {
            U[0] = F[0] * inv;
            U[1] = F[1] * inv;
            U[2] = F[2] * inv;
            tc[0] = F[3] * inv - U[0] * U[0];
            tc[1] = F[4] * inv - U[0] * U[1];
            tc[2] = F[5] * inv - U[0] * U[2];
            tc[3] = F[6] * inv - U[1] * U[1];
            tc[4] = F[7] * inv - U[1] * U[2];
            tc[5] = F[8] * inv - U[2] * U[2];
        }
    }
    double *ub = ubar + b * 3 * P;
    T *tb = tau + b * 6 * P;
    double o0, o1, o2;
    if (G.mode == 1) {
        const double sl = G.trig[i], cl = G.trig[Nx + i];
        const double sp = G.trig[2 * Nx + j], cp = G.trig[2 * Nx + Ny + j];
        const double ee[3] = {-sl, cl, 0.0}, en[3] = {-sp * cl, -sp * sl, cp}, er[3] = {cp * cl, cp * sl, sp};
        o0 = ee[0] * U[0] + ee[1] * U[1];
        o1 = en[0] * U[0] + en[1] * U[1] + en[2] * U[2];
        o2 = er[0] * U[0] + er[1] * U[1] + er[2] * U[2];
        const double T3[3][3] = {{tc[0], tc[1], tc[2]}, {tc[1], tc[3], tc[4]}, {tc[2], tc[4], tc[5]}};
        double Te[3], Tn[3];
#pragma unroll
        for (int rr = 0; rr < 3; rr++) {
            Te[rr] = T3[rr][0] * ee[0] + T3[rr][1] * ee[1] + T3[rr][2] * ee[2];
            Tn[rr] = T3[rr][0] * en[0] + T3[rr][1] * en[1] + T3[rr][2] * en[2];
        }
        tb[t] = (T)(ee[0] * Te[0] + ee[1] * Te[1] + ee[2] * Te[2]);
        tb[P + t] = (T)(ee[0] * Tn[0] + ee[1] * Tn[1] + ee[2] * Tn[2]);
        tb[2 * P + t] = (T)(en[0] * Tn[0] + en[1] * Tn[1] + en[2] * Tn[2]);
    } else if (G.mode == 0) {
        o0 = U[0];
        o1 = U[1];
        o2 = 0.0;
        tb[t] = (T)tc[0];
        tb[P + t] = (T)tc[1];
        tb[2 * P + t] = (T)tc[3];
    } else {
        o0 = U[0];
        o1 = U[1];
        o2 = U[2];
#pragma unroll
        for (int q6 = 0; q6 < 6; q6++) tb[q6 * P + t] = (T)tc[q6];
    }
    // the strain reads planes 0, 1 (and 2 in 3-D); the filtered velocities leave in T when asked for
    ub[t] = o0;
    ub[P + t] = o1;
    if (G.mode == 2) ub[2 * P + t] = o2;
    if (uo) uo[b * P + t] = (T)o0;
    if (vo) vo[b * P + t] = (T)o1;
    if (wo) wo[b * P + t] = (T)o2;
}

// The same with the normalization as a denominator: active points with den > 1e-15 (Filtering.jl:
// 1633) get F/den.
template <typename T>
__device__ __forceinline__ void pi_point_epilogue(const GeomArgs &G, int64_t t, int64_t i, int64_t j, const double *F,
                                                  double den, int64_t b, double *__restrict__ ubar, T *__restrict__ tau,
                                                  T *__restrict__ uo, T *__restrict__ vo, T *__restrict__ wo) {
    const bool ok = G.mask[t] != 0 && den > 1e-15;
    pi_point_epilogue_inv<T>(G, t, i, j, F, ok, ok ? 1.0 / den : 0.0, b, ubar, tau, uo, vo, wo);
}

}  // namespace cgef
