// asph_adjoint.h - derivative building blocks of the reverse (adjoint) sweep of one KDK step.
//
// Shared by the gfx950 reverse kernels (asph_adjoint.hip) and the FP64 host adjoint of the test
// oracle; templated on the real type R (float on the device, double on the host).
//
// Pointwise stages (EOS + microphysics + surface flux, the momentum epilogue, the density ratio)
// are written once for a generic number type and differentiated with a one-lane dual number:
// a vector-Jacobian product over n inputs takes n evaluations, each seeding one input. The pair
// terms of the density and momentum sums are differentiated by hand (the formulas follow the
// forward kernels' expressions, asph_h100.cu:306-333 and 444-491). Clamps contribute a zero
// derivative when active, branches follow the forward's decision.
#pragma once

#include <math.h>

#include "asph_internal.h"

namespace asph {
namespace adj {

// ---- scalar functions per real type ------------------------------------------------------------
template <class R> struct Fn;
template <> struct Fn<float> {
    static ASPH_DH float exp(float x) { return asphm::fexp(x); }
    static ASPH_DH float pow(float x, float y) { return asphm::fpow(x, y); }
    static ASPH_DH float sqrt(float x) { return sqrtf(x); }
    static ASPH_DH float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    // reciprocal for the reverse's pair terms, which need no correctly rounded division: the
    // hardware reciprocal (1 ulp) on the device, a division on the host
    static ASPH_DH float rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rcpf(x);
#else
        return 1.0f / x;
#endif
    }
};
template <> struct Fn<double> {
    static ASPH_DH double exp(double x) { return ::exp(x); }
    static ASPH_DH double pow(double x, double y) { return ::pow(x, y); }
    static ASPH_DH double sqrt(double x) { return ::sqrt(x); }
    static ASPH_DH double fma(double a, double b, double c) { return ::fma(a, b, c); }
    static ASPH_DH double rcp(double x) { return 1.0 / x; }
};

// ---- one-lane dual number ----------------------------------------------------------------------
template <class R> struct Dual {
    R v, d;
    ASPH_DH Dual() : v(0), d(0) {}
    ASPH_DH Dual(R x) : v(x), d(0) {}
    ASPH_DH Dual(R x, R dx) : v(x), d(dx) {}
};
template <class R> ASPH_DH Dual<R> operator+(Dual<R> a, Dual<R> b) { return {a.v + b.v, a.d + b.d}; }
template <class R> ASPH_DH Dual<R> operator-(Dual<R> a, Dual<R> b) { return {a.v - b.v, a.d - b.d}; }
template <class R> ASPH_DH Dual<R> operator-(Dual<R> a) { return {-a.v, -a.d}; }
template <class R> ASPH_DH Dual<R> operator*(Dual<R> a, Dual<R> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class R> ASPH_DH Dual<R> operator/(Dual<R> a, Dual<R> b) {
    R q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
template <class R> ASPH_DH bool operator<(Dual<R> a, Dual<R> b) { return a.v < b.v; }
template <class R> ASPH_DH bool operator>(Dual<R> a, Dual<R> b) { return a.v > b.v; }
template <class R> ASPH_DH Dual<R>& operator+=(Dual<R>& a, Dual<R> b) { a = a + b; return a; }
template <class R> ASPH_DH Dual<R>& operator-=(Dual<R>& a, Dual<R> b) { a = a - b; return a; }

// generic operations on R or Dual<R>
template <class R> ASPH_DH R gval(R x) { return x; }
template <class R> ASPH_DH R gval(Dual<R> x) { return x.v; }
template <class R> ASPH_DH R gexp(R x) { return Fn<R>::exp(x); }
template <class R> ASPH_DH Dual<R> gexp(Dual<R> x) {
    R e = Fn<R>::exp(x.v);
    return {e, e * x.d};
}
template <class R> ASPH_DH R gpow(R x, R y) { return Fn<R>::pow(x, y); }
template <class R> ASPH_DH Dual<R> gpow(Dual<R> x, R y) {
    R p = Fn<R>::pow(x.v, y);
    return {p, x.v != R(0) ? ((y * p) / x.v) * x.d : R(0)};
}
template <class R> ASPH_DH R gmin(R a, R b) { return b < a ? b : a; }
template <class R> ASPH_DH R gmax(R a, R b) { return b > a ? b : a; }
template <class R> ASPH_DH Dual<R> gmin(Dual<R> a, Dual<R> b) { return b.v < a.v ? b : a; }
template <class R> ASPH_DH Dual<R> gmax(Dual<R> a, Dual<R> b) { return b.v > a.v ? b : a; }

template <class T> struct RealOf { using type = T; };
template <class R> struct RealOf<Dual<R>> { using type = R; };

// ---- reference state and thermodynamics (asph_internal.h, generic) ------------------------------
template <class T> ASPH_DH T g_theta_ref(const DevParams& P, T z) {
    using R = typename RealOf<T>::type;
    if (P.n2 > 0.0f) return T(R(P.th0)) * gexp((T(R(P.n2)) * z) / T(R(GRAV)));
    return T(R(P.th0));
}
template <class T> ASPH_DH T g_exner_ref(const DevParams& P, T z) {
    using R = typename RealOf<T>::type;
    if (P.n2 > 0.0f)
        return T(R(1)) - T(R(P.exA)) * (T(R(1)) - gexp((T(R(-P.n2)) * z) / T(R(GRAV))));
    return T(R(1)) - (T(R(GRAV)) * z) / T(R(CP) * R(P.th0));
}
template <class T> ASPH_DH T g_pres_of_exner(const DevParams& P, T ex) {
    using R = typename RealOf<T>::type;
    return T(R(P0)) * gpow(ex, R(P.cp_rd));
}
template <class T> ASPH_DH T g_rho_ref(const DevParams& P, T z) {
    using R = typename RealOf<T>::type;
    T ex = g_exner_ref(P, z);
    return g_pres_of_exner(P, ex) / ((T(R(R_D)) * g_theta_ref(P, z)) * ex);
}
template <class T> ASPH_DH T g_esat(T Tk) {
    using R = typename RealOf<T>::type;
    T Tc = Tk - T(R(273.15f));
    return T(R(611.2f)) * gexp((T(R(17.67f)) * Tc) / (Tc + T(R(243.5f))));
}
template <class T> ASPH_DH T g_qsat(T Tk, T p) {
    using R = typename RealOf<T>::type;
    T es = gmin(g_esat(Tk), T(R(0.5f)) * p);
    T den = p - T(R(0.378f)) * es;
    return gval(den) > R(1) ? (T(R(0.622f)) * es) / den : T(R(0));
}
template <class T> ASPH_DH T g_dqsdT(T Tk, T es, T p) {
    using R = typename RealOf<T>::type;
    T Tc = Tk - T(R(273.15f));
    T d = Tc + T(R(243.5f));
    T desdT = (es * T(R(17.67f * 243.5f))) / (d * d);
    T den = p - T(R(0.378f)) * es;
    return ((T(R(0.622f)) * p) * desdT) / (den * den);
}
template <class T> ASPH_DH T g_rh_interp(const DevParams& P, T z) {
    using R = typename RealOf<T>::type;
    const int n = P.rh_n;
    if (n <= 0) return T(R(0));
    const R zv = gval(z);
    if (zv <= R(P.rh_z[0])) return T(R(P.rh_v[0]));
    R zl = R(P.rh_z[0]), vl = R(P.rh_v[0]);
#pragma unroll
    for (int k = 1; k < ASPH_RH_KNOTS; ++k)
        if (k == n - 1) { zl = R(P.rh_z[k]); vl = R(P.rh_v[k]); }
    if (zv >= zl) return T(vl);
    R za = R(P.rh_z[0]), va = R(P.rh_v[0]), zb = R(P.rh_z[1]), vb = R(P.rh_v[1]);
#pragma unroll
    for (int k = 1; k + 1 < ASPH_RH_KNOTS; ++k)
        if (k + 1 < n && zv >= R(P.rh_z[k])) {
            za = R(P.rh_z[k]); va = R(P.rh_v[k]);
            zb = R(P.rh_z[k + 1]); vb = R(P.rh_v[k + 1]);
        }
    T w = (z - T(za)) / T(zb - za);
    return T(va) + T(vb - va) * w;
}
template <class T> ASPH_DH T g_qv_ref(const DevParams& P, T z) {
    using R = typename RealOf<T>::type;
    if (P.qv_mode == ASPH_QV_REF_EXP) {
        T q = T(R(P.qv0)) * gexp(-z / T(R(P.qv_hs)));
        T d = gmax(z - T(R(P.qv_zt)), T(R(0)));
        return q * gexp(-d / T(R(P.qv_ht)));
    }
    if (P.qv_mode == ASPH_QV_RH_TABLE) {
        T ex = g_exner_ref(P, z);
        T Tk = g_theta_ref(P, z) * ex;
        T p = g_pres_of_exner(P, ex);
        return g_rh_interp(P, z) * g_qsat(Tk, p);
    }
    return T(R(0));
}

// ---- EOS + warm-rain microphysics of one particle (density epilogue) ---------------------------
// Inputs (in order): theta, qv, qc at the start of the step, the new density, z, the rain of the
// particle's cell at the start of the step [particle masses], and the autoconversion threshold.
// Outputs: theta, qv, qc after the stage, p', and the net rain deposit dq [kg/kg] (autoconversion
// + accretion - evaporation). `flux` applies the bulk surface flux first (asph_h100.cu:606-626).
constexpr int EOS_IN = 7;
constexpr int EOS_OUT = 5;

template <class R> struct EosConst {
    R dt, auto_f, Ts, qs_sfc;
    bool flux;
};

template <class T, class R>
ASPH_DH void eos_stage(const DevParams& P, const EosConst<R>& K, const T in[EOS_IN], T out[EOS_OUT]) {
    T th = in[0], qv = in[1], qc = in[2];
    const T rnew = in[3], z = in[4], rf = in[5], qc0 = in[6];
    const T ex = g_exner_ref(P, z);
    const T pref = g_pres_of_exner(P, ex);
    const T rref = pref / ((T(R(R_D)) * g_theta_ref(P, z)) * ex);
    if (K.flux) {
        T Pi = gpow(gmax(pref, T(R(1))) / T(R(P0)), R(P.rd_cp));
        T Ta = th * Pi;
        th = th + ((T(R(P.c_h) * R(P.v_bg)) * (T(K.Ts) - Ta)) * T(K.dt)) / T(R(P.dz));
        qv = qv + ((T((R(P.beta_moist) * R(P.c_e)) * R(P.v_bg)) * (T(K.qs_sfc) - qv)) * T(K.dt)) / T(R(P.dz));
        th = gmin(gmax(th, T(R(100))), T(R(P.theta_cap)));
        qv = gmin(gmax(qv, T(R(0))), T(R(P.qv_cap)));
    }
    const T r = gmax(rnew, T(R(1e-3f)));
    T dd = r - rref;
    dd = gmin(gmax(dd, T(R(-0.3f)) * rref), T(R(0.3f)) * rref);
    const T ppr = T(R(P.cs2)) * dd;
    const T pfl = gmax(pref + ppr, T(R(1)));
    const T Pi = gpow(pfl / T(R(P0)), R(P.rd_cp));
    const T LvCp = T(R(L_V) / R(CP));
    T Tk = th * Pi;
    const T qs = g_qsat(Tk, pfl);
    if (qv > qs || gval(qc) > R(0)) {
        T dqc = T(R(0));
        for (int it = 0; it < 3; ++it) {
            T Tn = Tk + LvCp * dqc;
            T esn = g_esat(Tn);
            T qsn = g_qsat(Tn, pfl);
            T f = (qv - dqc) - qsn;
            T fp = T(R(-1)) - g_dqsdT(Tn, esn, pfl) * LvCp;
            dqc = dqc - f / fp;
            dqc = gmin(dqc, qv);
            dqc = gmax(dqc, -qc);
        }
        qv = qv - dqc;
        qc = qc + dqc;
        th = th + (T(R(L_V)) * dqc) / (T(R(CP)) * Pi);
        Tk = Tk + LvCp * dqc;
    }
    T dq = T(R(0));
    if ((P.flags & ASPH_F_AUTOCONV) && qc > qc0) {
        T dqr = (qc - qc0) * T(K.auto_f);
        qc = qc - dqr;
        dq = dq + dqr;
    }
    if (P.flags & (ASPH_F_ACCRETION | ASPH_F_RAIN_EVAP)) {
        T qr = (rf * T(R(P.mm))) / r;
        if ((P.flags & ASPH_F_ACCRETION) && gval(qc) > R(0) && gval(qr) > R(0)) {
            T qcn = qc / (T(R(1)) + T(R(P.k_accr) * K.dt) * gpow(qr, R(0.875f)));
            dq = dq + (qc - qcn);
            qc = qcn;
        }
        if ((P.flags & ASPH_F_RAIN_EVAP) && gval(qr) > R(0)) {
            T qs2 = g_qsat(Tk, pfl);
            if (qv < qs2) {
                T rc = T(R(0.001f)) * r;
                T rq = rc * qr;
                T C = T(R(1.6f)) + T(R(124.9f)) * gpow(rq, R(0.2046f));
                T E = (((T(R(1)) - qv / qs2) * C) * gpow(rq, R(0.525f))) /
                      ((T(R(5.4e5f)) + T(R(2.55e8f)) / (pfl * qs2)) * rc);
                T dqe = gmin(gmin(E * T(K.dt), qr), qs2 - qv);
                qv = qv + dqe;
                th = th - (T(R(L_V)) * dqe) / (T(R(CP)) * Pi);
                dq = dq - dqe;
            }
        }
    }
    out[0] = gmin(gmax(th, T(R(100))), T(R(P.theta_cap)));
    out[1] = gmin(gmax(qv, T(R(0))), T(R(P.qv_cap)));
    out[2] = gmax(qc, T(R(0)));
    out[3] = ppr;
    out[4] = dq;
}

// ---- momentum epilogue of one particle (asph_h100.cu:497-510, theta update, kick, sponge) -------
// Inputs: the pair sums times mm (a_x, a_y, a_z, av_x, av_y, av_z), v after the drift (3), z,
// theta, qv, qc after the EOS stage, p'. Outputs: acc (3), theta after the update, v after the
// second half kick and the sponge (3). mode 0 (force evaluation): acc only.
constexpr int MOM_IN = 14;
constexpr int MOM_OUT = 7;

template <class T, class R>
ASPH_DH void mom_stage(const DevParams& P, R dt, int mode, const T in[MOM_IN], T out[MOM_OUT]) {
    const T z = in[9], th = in[10], qv = in[11], qc = in[12], pr = in[13];
    const T thv = th * (T(R(1)) + T(R(0.608f)) * qv);
    const T thvr = g_theta_ref(P, z) * (T(R(1)) + T(R(0.608f)) * g_qv_ref(P, z));
    const T ax = in[0], ay = in[1];
    const T az = in[2] + T(R(GRAV)) * ((thv - thvr) / thvr - qc);
    out[0] = ax; out[1] = ay; out[2] = az;
    if (mode == 0) {
        out[3] = th; out[4] = in[6]; out[5] = in[7]; out[6] = in[8];
        return;
    }
    const T ex = g_exner_ref(P, z);
    const T Pi = gpow(gmax(g_pres_of_exner(P, ex) + pr, T(R(1))) / T(R(P0)), R(P.rd_cp));
    const T vdot = in[8] * in[5] + (in[7] * in[4] + in[6] * in[3]);
    T dth = -vdot / (T(R(CP)) * Pi);
    dth = gmin(gmax(dth, T(R(-P.dthdt_cap))), T(R(P.dthdt_cap)));
    out[3] = gmin(gmax(th + dth * T(dt), T(R(100))), T(R(P.theta_cap)));
    const T hdt = T(R(0.5f) * dt);
    T vx = in[6] + ax * hdt, vy = in[7] + ay * hdt, vz = in[8] + az * hdt;
    if ((P.flags & ASPH_F_SPONGE) && gval(z) >= R(P.sponge_zs)) {
        T tt = (z - T(R(P.sponge_zs))) / T(R(P.sponge_depth));
        T sg = (T(R(P.sponge_sigma)) * tt) * tt;
        vz = vz * (T(R(1)) / (T(R(1)) + sg * T(dt)));
    }
    out[4] = vx; out[5] = vy; out[6] = vz;
}

// ---- density ratio (asph_h100.cu:331-332) --------------------------------------------------------
// Inputs: S1, S2, z. Output: the clamped density.
template <class T, class R> ASPH_DH T dens_stage(const DevParams& P, T S1, T S2, T z) {
    const T rref = g_rho_ref(P, z);
    const T rr = gval(S2) > R(1e-20f) ? S1 / S2 : rref;
    return gmin(gmax(rr, T(R(0.5f)) * rref), T(R(2.0f)) * rref);
}

// ---- vector-Jacobian products of the pointwise stages --------------------------------------------
template <class R>
ASPH_DH void eos_vjp(const DevParams& P, const EosConst<R>& K, const R in[EOS_IN], const R obar[EOS_OUT],
                     R ibar[EOS_IN]) {
    for (int k = 0; k < EOS_IN; ++k) {
        Dual<R> x[EOS_IN], o[EOS_OUT];
        for (int j = 0; j < EOS_IN; ++j) x[j] = Dual<R>(in[j], j == k ? R(1) : R(0));
        eos_stage<Dual<R>, R>(P, K, x, o);
        R s = R(0);
        for (int m = 0; m < EOS_OUT; ++m) s += o[m].d * obar[m];
        ibar[k] = s;
    }
}

template <class R>
ASPH_DH void mom_vjp(const DevParams& P, R dt, int mode, const R in[MOM_IN], const R obar[MOM_OUT],
                     R ibar[MOM_IN]) {
    for (int k = 0; k < MOM_IN; ++k) {
        Dual<R> x[MOM_IN], o[MOM_OUT];
        for (int j = 0; j < MOM_IN; ++j) x[j] = Dual<R>(in[j], j == k ? R(1) : R(0));
        mom_stage<Dual<R>, R>(P, dt, mode, x, o);
        R s = R(0);
        for (int m = 0; m < MOM_OUT; ++m) s += o[m].d * obar[m];
        ibar[k] = s;
    }
}

// The vector-Jacobian product of mom_stage in reverse mode: the stage's few nonlinear terms are
// differentiated by hand and its height-dependent reference profiles (theta_ref, qv_ref, the
// reference pressure) by one dual pass in z, instead of one forward pass per input. Clamps pass
// the derivative only where mom_stage's comparison selects the argument (the same tests).
template <class R>
ASPH_DH void mom_vjp_rev(const DevParams& P, R dt, int mode, const R in[MOM_IN], const R ob[MOM_OUT],
                         R ib[MOM_IN]) {
    for (int k = 0; k < MOM_IN; ++k) ib[k] = R(0);
    const R z = in[9], th = in[10], qv = in[11], qc = in[12], pr = in[13];
    const Dual<R> zd(z, R(1));
    const Dual<R> thvr_d = g_theta_ref(P, zd) * (Dual<R>(R(1)) + Dual<R>(R(0.608f)) * g_qv_ref(P, zd));
    const R thvr = thvr_d.v;
    const R thv = th * (R(1) + R(0.608f) * qv);
    R ab[3] = {ob[0], ob[1], ob[2]};   // adjoints of ax, ay, az
    R thb = R(0), zb = R(0);
    if (mode == 0) {
        thb = ob[3];
        ib[6] = ob[4]; ib[7] = ob[5]; ib[8] = ob[6];
    } else {
        const R az = in[2] + R(GRAV) * ((thv - thvr) / thvr - qc);
        const Dual<R> pz_d = g_pres_of_exner(P, g_exner_ref(P, zd));
        const R u_raw = pz_d.v + pr;
        const bool u_on = !(R(1) > u_raw);   // gmax(p + p', 1) selects p + p'
        const R u = u_on ? u_raw : R(1);
        const R Pi = Fn<R>::pow(u / R(P0), R(P.rd_cp));
        const R vdot = in[8] * in[5] + (in[7] * in[4] + in[6] * in[3]);
        const R dth_raw = -vdot / (R(CP) * Pi);
        const R cap = R(P.dthdt_cap);
        const bool d_on = !(-cap > dth_raw) && !(cap < dth_raw);
        const R dth = d_on ? dth_raw : (-cap > dth_raw ? -cap : cap);
        const R th2 = th + dth * dt;
        const bool t_on = !(R(100) > th2) && !(R(P.theta_cap) < th2);
        const R hdt = R(0.5f) * dt;
        // velocities (second half kick, sponge on vz)
        R vzb = ob[6];
        if ((P.flags & ASPH_F_SPONGE) && z >= R(P.sponge_zs)) {
            const R tt = (z - R(P.sponge_zs)) / R(P.sponge_depth);
            const R sg = (R(P.sponge_sigma) * tt) * tt;
            const R f = R(1) / (R(1) + sg * dt);
            const R vz0 = in[8] + az * hdt;
            const R sgb = ob[6] * vz0 * (-(f * f) * dt);
            zb += sgb * (R(2) * R(P.sponge_sigma) * tt) / R(P.sponge_depth);
            vzb = ob[6] * f;
        }
        ib[6] += ob[4]; ib[7] += ob[5]; ib[8] += vzb;
        ab[0] += ob[4] * hdt; ab[1] += ob[5] * hdt; ab[2] += vzb * hdt;
        // theta update
        const R th2b = t_on ? ob[3] : R(0);
        thb += th2b;
        const R dthb = d_on ? th2b * dt : R(0);
        const R vdotb = -dthb / (R(CP) * Pi);
        const R Pib = dthb * vdot / ((R(CP) * Pi) * Pi);
        ib[5] += vdotb * in[8]; ib[8] += vdotb * in[5];
        ib[4] += vdotb * in[7]; ib[7] += vdotb * in[4];
        ib[3] += vdotb * in[6]; ib[6] += vdotb * in[3];
        if (u_on) {
            const R ub = Pib * (R(P.rd_cp) * Pi / u);
            ib[13] += ub;
            zb += ub * pz_d.d;
        }
    }
    // buoyancy of az: GRAV ((thv - thvr) / thvr - qc)
    ib[0] += ab[0]; ib[1] += ab[1]; ib[2] += ab[2];
    const R bb = ab[2] * R(GRAV);
    const R thvb = bb / thvr;
    zb += (-bb * thv / (thvr * thvr)) * thvr_d.d;
    ib[12] += -bb;
    thb += thvb * (R(1) + R(0.608f) * qv);
    ib[11] += thvb * th * R(0.608f);
    ib[10] += thb;
    ib[9] += zb;
}

// d(rho)/d(S1, S2, z) times rho_bar.
template <class R> ASPH_DH void dens_vjp(const DevParams& P, R S1, R S2, R z, R rbar, R& s1bar, R& s2bar, R& zbar) {
    const R in[3] = {S1, S2, z};
    R ib[3];
    for (int k = 0; k < 3; ++k) {
        Dual<R> x[3];
        for (int j = 0; j < 3; ++j) x[j] = Dual<R>(in[j], j == k ? R(1) : R(0));
        Dual<R> o = dens_stage<Dual<R>, R>(P, x[0], x[1], x[2]);
        ib[k] = o.d * rbar;
    }
    s1bar = ib[0]; s2bar = ib[1]; zbar = ib[2];
}

// ---- pair terms --------------------------------------------------------------------------------
// Density kernel value W(r) = f(q) and dW/dr for r = x_owner - x_neighbour, given as the
// separation in cells (dxm, dym, dzm) = (x_o / dx - x_n / dx, ...) with the minimum image applied
// by the caller, as the forward takes it. Returns false outside the support (the forward's test
// r2 < 4h^2, same arithmetic).
template <class R>
ASPH_DH bool dens_pair(const DevParams& P, R dxm, R dym, R dzm, R& W, R gw[3]) {
    const R ix = R(P.inv_dx), iy = R(P.inv_dy), iz = R(P.inv_dz);
    const R r2 = Fn<R>::fma(dzm, dzm, Fn<R>::fma(dym, dym, dxm * dxm));
    if (!(r2 < R(P.lim))) return false;
    const R q = Fn<R>::sqrt(r2) * R(P.inv_h);
    R kq;   // f'(q) / q
    if (q < R(1)) {
        W = Fn<R>::fma(q * q, Fn<R>::fma(R(0.75f), q, R(-1.5f)), R(1));
        kq = R(2.25f) * q - R(3);
    } else {
        const R t = R(2) - q;
        W = (R(0.25f) * t) * (t * t);
        kq = -(R(0.75f) * t * t) * Fn<R>::rcp(q);
    }
    const R f = kq * R(P.inv_h) * R(P.inv_h);
    gw[0] = f * dxm * ix; gw[1] = f * dym * iy; gw[2] = f * dzm * iz;
    return true;
}

// Momentum pair (asph_h100.cu:460-491): owner o, neighbour n, r = x_o - x_n, given as the
// separation in cells (sx, sy, sz) as the forward takes it (r = s times (dx, dy, dz)). Given the
// owner's adjoints of its pressure/viscous sum (Ab = d/d sum(-br g)) and of its viscous-only sum
// (AVb = d/d sum(-visc g)), returns the adjoints of r, v_o - v_n, rho_o, rho_n, p_o, p_n.
// False outside the support or for coincident particles (the forward's tests, same arithmetic).
template <class R> struct PairBar {
    R r[3], vd[3], rho_o, rho_n, p_o, p_n;
};

template <class R>
ASPH_DH bool mom_pair_vjp(const DevParams& P, R sx, R sy, R sz, const R vo[3], const R vn[3], R rho_o,
                          R rho_n, R p_o, R p_n, const R Ab[3], const R AVb[3], PairBar<R>& b) {
    const R invd[3] = {R(P.inv_dx), R(P.inv_dy), R(P.inv_dz)};
    const R dt3[3] = {sx, sy, sz};
    const R rx = sx * R(P.dx), ry = sy * R(P.dy), rz = sz * R(P.dz);
    const R rv[3] = {rx, ry, rz};
    const R r2t = Fn<R>::fma(dt3[2], dt3[2], Fn<R>::fma(dt3[1], dt3[1], dt3[0] * dt3[0]));
    if (!(r2t < R(P.lim)) || r2t < R(1e-12f)) return false;
    const R q = Fn<R>::sqrt(r2t) * R(P.inv_h);
    R kk, dk;   // kgrad(q) and its derivative
    if (q < R(1)) {
        kk = Fn<R>::fma(R(2.25f), q, R(-3.0f)) * R(P.kin);
        dk = R(2.25f) * R(P.kin);
    } else {
        const R t = R(2) - q;
        kk = ((R(P.kout) * t) * t) / q;
        dk = -(R(P.kout) * t * (q + R(2))) / (q * q);
    }
    R g[3];
    for (int a = 0; a < 3; ++a) g[a] = (kk * dt3[a]) * invd[a];
    const R inv_ro = R(1) / rho_o, inv_rn = R(1) / rho_n;
    const R dp = (p_n - p_o) * (inv_ro * inv_rn);
    const R vd[3] = {vo[0] - vn[0], vo[1] - vn[1], vo[2] - vn[2]};
    const R vdotr = Fn<R>::fma(vd[2], rz, Fn<R>::fma(vd[1], ry, vd[0] * rx));
    R visc = R(0), rp2 = R(0), rbar = R(0), D = R(0);
    const bool vis = vdotr < R(0);
    if (vis) {
        rp2 = Fn<R>::fma(rz, rz, Fn<R>::fma(ry, ry, rx * rx));
        rbar = R(0.5f) * (rho_o + rho_n);
        D = ((rp2 + R(P.eta2)) * (rbar * rbar)) * rho_n;
        visc = (R(P.visc_c) * vdotr) / D;
    }
    const R br = dp + visc;
    R G = R(0), Gv = R(0), Gb[3], S = R(0);
    for (int a = 0; a < 3; ++a) {
        G += Ab[a] * g[a];
        Gv += AVb[a] * g[a];
        Gb[a] = -br * Ab[a] - visc * AVb[a];   // d loss / d g
        S += Gb[a] * dt3[a] * invd[a];
    }
    const R c1 = ((dk * R(P.inv_h)) * R(P.inv_h)) / q * S;
    for (int a = 0; a < 3; ++a) b.r[a] = c1 * dt3[a] * invd[a] + kk * Gb[a] * invd[a] * invd[a];
    // dp: loss = -G * dp - (G + Gv) * visc
    b.p_n = -G * (inv_ro * inv_rn);
    b.p_o = -b.p_n;
    b.rho_o = G * dp * inv_ro;
    b.rho_n = G * dp * inv_rn;
    for (int a = 0; a < 3; ++a) b.vd[a] = R(0);
    if (vis) {
        const R Vb = -G - Gv;
        const R dv = Vb * (R(P.visc_c) / D);
        const R cr = Vb * (-visc / (rp2 + R(P.eta2))) * R(2);
        for (int a = 0; a < 3; ++a) {
            b.vd[a] = dv * rv[a];
            b.r[a] += dv * vd[a] + cr * rv[a];
        }
        b.rho_o += -Vb * visc / rbar;
        b.rho_n += Vb * (-visc / rbar - visc / rho_n);
    }
    return true;
}


// Both orientations of one momentum pair at once, as the reverse gather of particle i needs them:
// (i, j) of i's sum with i's seeds (Ai, AVi) and (j, i) of j's sum with j's seeds (Aj, AVj),
// r = x_i - x_j. The geometry (kernel gradient, density and pressure factors, the viscous
// denominator's common part) is evaluated once; the terms are those of mom_pair_vjp, with the
// second orientation's r, v_o - v_n and kernel gradient negated. A zero seed drops its
// orientation (the terms are linear in the seeds). Adds the adjoints of x_i, v_i, rho_i and p_i.
// r = x_i - x_j is given in cells, as for mom_pair_vjp. False outside the support or for
// coincident particles.
template <class R>
ASPH_DH bool mom_pair_vjp2(const DevParams& P, R sx, R sy, R sz, const R vi[3], const R vj[3], R rho_i, R rho_j,
                           R p_i, R p_j, const R Ai[3], const R AVi[3], const R Aj[3], const R AVj[3], R xb[3],
                           R vb[3], R& rb, R& pb) {
    const R invd[3] = {R(P.inv_dx), R(P.inv_dy), R(P.inv_dz)};
    const R dt3[3] = {sx, sy, sz};
    const R rx = sx * R(P.dx), ry = sy * R(P.dy), rz = sz * R(P.dz);
    const R rv[3] = {rx, ry, rz};
    const R r2t = Fn<R>::fma(dt3[2], dt3[2], Fn<R>::fma(dt3[1], dt3[1], dt3[0] * dt3[0]));
    if (!(r2t < R(P.lim)) || r2t < R(1e-12f)) return false;
    const R q = Fn<R>::sqrt(r2t) * R(P.inv_h);
    const R iq = Fn<R>::rcp(q);
    R kk, dk;
    if (q < R(1)) {
        kk = Fn<R>::fma(R(2.25f), q, R(-3.0f)) * R(P.kin);
        dk = R(2.25f) * R(P.kin);
    } else {
        const R t = R(2) - q;
        kk = ((R(P.kout) * t) * t) * iq;
        dk = -(R(P.kout) * t * (q + R(2))) * (iq * iq);
    }
    const R c1f = ((dk * R(P.inv_h)) * R(P.inv_h)) * iq;
    R u[3], g[3];
    for (int a = 0; a < 3; ++a) {
        u[a] = dt3[a] * invd[a];
        g[a] = kk * u[a];
    }
    const R inv_i = Fn<R>::rcp(rho_i), inv_j = Fn<R>::rcp(rho_j), inv_ij = inv_i * inv_j;
    const R dp = (p_j - p_i) * inv_ij;   // (i, j); (j, i) has -dp
    const R vd[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
    const R vdotr = Fn<R>::fma(vd[2], rz, Fn<R>::fma(vd[1], ry, vd[0] * rx));
    const bool vis = vdotr < R(0);
    // viscous denominators D = (r^2 + eta^2) rbar^2 rho_n: of (i, j) with rho_j, of (j, i) with rho_i
    R visc_i = R(0), visc_j = R(0), iDi = R(0), iDj = R(0), irpe = R(0), irbar = R(0);
    if (vis) {
        const R rp2 = Fn<R>::fma(rz, rz, Fn<R>::fma(ry, ry, rx * rx));
        const R rbar = R(0.5f) * (rho_i + rho_j);
        const R rpe = rp2 + R(P.eta2);
        irbar = Fn<R>::rcp(rbar);
        irpe = Fn<R>::rcp(rpe);
        const R ibase = irpe * (irbar * irbar);
        iDj = ibase * inv_j;
        iDi = ibase * inv_i;
        visc_i = (R(P.visc_c) * vdotr) * iDj;
        visc_j = (R(P.visc_c) * vdotr) * iDi;
    }
    // (i, j): owner i
    {
        const R br = dp + visc_i;
        R G = R(0), Gv = R(0), S = R(0), Gb[3];
        for (int a = 0; a < 3; ++a) {
            G += Ai[a] * g[a];
            Gv += AVi[a] * g[a];
            Gb[a] = -br * Ai[a] - visc_i * AVi[a];
            S += Gb[a] * u[a];
        }
        for (int a = 0; a < 3; ++a) xb[a] += c1f * S * u[a] + kk * Gb[a] * invd[a] * invd[a];
        pb += G * inv_ij;
        rb += G * dp * inv_i;
        if (vis) {
            const R Vb = -G - Gv;
            const R dv = Vb * (R(P.visc_c) * iDj);
            const R cr = Vb * (-visc_i * irpe) * R(2);
            for (int a = 0; a < 3; ++a) {
                vb[a] += dv * rv[a];
                xb[a] += dv * vd[a] + cr * rv[a];
            }
            rb += -Vb * visc_i * irbar;
        }
    }
    // (j, i): owner j, r' = -r, g' = -g, v_o - v_n = -vd; the adjoint of x_i is minus that of r'
    {
        const R br = -dp + visc_j;
        R G = R(0), Gv = R(0), S = R(0), Gb[3];
        for (int a = 0; a < 3; ++a) {
            G -= Aj[a] * g[a];
            Gv -= AVj[a] * g[a];
            Gb[a] = -br * Aj[a] - visc_j * AVj[a];
            S += Gb[a] * u[a];   // S' = -S
        }
        for (int a = 0; a < 3; ++a) xb[a] -= c1f * S * u[a] + kk * Gb[a] * invd[a] * invd[a];
        pb -= G * inv_ij;        // p_n of (j, i) is p_i
        rb += G * (-dp) * inv_i; // rho_n of (j, i) is rho_i
        if (vis) {
            const R Vb = -G - Gv;
            const R dv = Vb * (R(P.visc_c) * iDi);
            const R cr = Vb * (-visc_j * irpe) * R(2);
            for (int a = 0; a < 3; ++a) {
                vb[a] += dv * rv[a];                   // -(dv * r')
                xb[a] += dv * vd[a] + cr * rv[a];      // -(dv * (-vd) + cr * r')
            }
            rb += Vb * (-visc_j * irbar - visc_j * inv_i);
        }
    }
    return true;
}

}  // namespace adj
}  // namespace asph
