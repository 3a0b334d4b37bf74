// Hyperelastic PK1 stress for the three laws of src/material/*.cpp, written
// in terms of the displacement gradient H = F - I so that FP32 evaluation
// keeps its relative accuracy at small strain:
//   * J - 1 = tr H + (tr^2 H - tr H^2)/2 + det H, always in FP64;
//   * E = C - I = H + H^T + H^T H (twice the Green strain), from H directly;
//   * the isochoric terms use F^-T dev(.) forms with no O(1) cancellation:
//       2F - 2/3 I1 F^-T             = 2 F^-T dev(E)
//       2F(I1 I - C) - 4/3 I2 F^-T   = 2 F^-T dev((1 + tr E) E - E^2)
//   so P = 2 c1 J^-2/3 F^-T dev(E)                         [I1bar term]
//        + 2 C01 J^-4/3 F^-T dev((1 + tr E) E - E^2)       [I2bar term, MR]
//        + (2/D1)(J - 1) cof F                              [volumetric]
// with c1 = C10 (NH, MR) or C10 + 2 C20 t + 3 C30 t^2, t = I1bar - 3 (Yeoh),
// which equals NeoHookean.cpp:4-20, MooneyRivlin.cpp:4-24 and Yeoh.cpp:4-20
// (dJ/dF = J F^-T = cof F, I1bar = I1 / I3^(1/3), I2bar = I2 / I3^(2/3),
// I3 = J^2, so J^-2/3 stands for |J|^-2/3 throughout, inverted elements
// included).
// Strain energies (no counterpart in the reference, which never evaluates
// one): W = c(I1bar-3)[+C01(I2bar-3)] + (J-1)^2/D1, Yeoh sum_i Ci0(I1bar-3)^i.
#pragma once
#include <type_traits>

#ifdef __HIPCC__
#define EFEM_HD __host__ __device__ __forceinline__
#else
#define EFEM_HD inline
#endif

#include <cmath>

namespace efem {

enum { LAW_NH = 0, LAW_MR = 1, LAW_YEOH = 2 };

// calibrated "t=2.0" coefficient of the approximate Jacobian
// (NeoHookean.cpp:41, MooneyRivlin.cpp:48, Yeoh.cpp:47)
EFEM_HD double elasticityCoeff(int law, const double* p)
{
    // evaluated the same way on the host and the device (the device
    // assembly of the approximate Jacobian reproduces the host's bit for bit)
#pragma clang fp contract(off)
    if (law == LAW_NH) return 2.94001 * p[0] + 2.0 / p[3];
    if (law == LAW_MR) return 3.818 * p[1] + 2.94001 * p[0] + 2.0 / p[3];
    return 2.94001 * p[0] + 3.51194 * p[1] + 3.3154 * p[2] + 2.0 / p[3];
}

// ---- forward-mode dual numbers ---------------------------------------------
// v + d eps with eps^2 = 0: running the stress code on Dual<T> gives the
// directional derivative dP = (dP/dF)[dF] exactly (the material tangent
// applied to dF), which the tangent/Hessian-vector kernels use.
template <typename S>
struct Dual {
    S v, d;
    EFEM_HD Dual() : v(0), d(0) {}
    EFEM_HD Dual(S a) : v(a), d(0) {}
    EFEM_HD Dual(S a, S b) : v(a), d(b) {}
    template <typename S2>
    EFEM_HD explicit Dual(const Dual<S2>& o) : v(static_cast<S>(o.v)), d(static_cast<S>(o.d)) {}
    EFEM_HD Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    EFEM_HD Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
};
template <typename S> EFEM_HD Dual<S> operator-(const Dual<S>& a) { return Dual<S>(-a.v, -a.d); }
template <typename S> EFEM_HD Dual<S> operator+(const Dual<S>& a, const Dual<S>& b) { return Dual<S>(a.v + b.v, a.d + b.d); }
template <typename S> EFEM_HD Dual<S> operator-(const Dual<S>& a, const Dual<S>& b) { return Dual<S>(a.v - b.v, a.d - b.d); }
template <typename S> EFEM_HD Dual<S> operator*(const Dual<S>& a, const Dual<S>& b) { return Dual<S>(a.v * b.v, a.v * b.d + a.d * b.v); }
template <typename S> EFEM_HD Dual<S> operator/(const Dual<S>& a, const Dual<S>& b)
{
    const S q = a.v / b.v;
    return Dual<S>(q, (a.d - q * b.d) / b.v);
}
template <typename S> EFEM_HD Dual<S> operator+(const Dual<S>& a, S b) { return Dual<S>(a.v + b, a.d); }
template <typename S> EFEM_HD Dual<S> operator+(S a, const Dual<S>& b) { return Dual<S>(a + b.v, b.d); }
template <typename S> EFEM_HD Dual<S> operator-(const Dual<S>& a, S b) { return Dual<S>(a.v - b, a.d); }
template <typename S> EFEM_HD Dual<S> operator-(S a, const Dual<S>& b) { return Dual<S>(a - b.v, -b.d); }
template <typename S> EFEM_HD Dual<S> operator*(const Dual<S>& a, S b) { return Dual<S>(a.v * b, a.d * b); }
template <typename S> EFEM_HD Dual<S> operator*(S a, const Dual<S>& b) { return Dual<S>(a * b.v, a * b.d); }
template <typename S> EFEM_HD Dual<S> operator/(const Dual<S>& a, S b) { return Dual<S>(a.v / b, a.d / b); }
template <typename S> EFEM_HD Dual<S> operator/(S a, const Dual<S>& b)
{
    const S q = a / b.v;
    return Dual<S>(q, -q * b.d / b.v);
}

EFEM_HD double expm1_(double x) { return std::expm1(x); }
EFEM_HD double log1p_(double x) { return std::log1p(x); }
EFEM_HD double sqrt_(double x) { return std::sqrt(x); }
template <typename S> EFEM_HD Dual<S> expm1_(const Dual<S>& x) { return Dual<S>(expm1_(x.v), (expm1_(x.v) + S(1)) * x.d); }
template <typename S> EFEM_HD Dual<S> log1p_(const Dual<S>& x) { return Dual<S>(log1p_(x.v), x.d / (S(1) + x.v)); }
template <typename S> EFEM_HD Dual<S> sqrt_(const Dual<S>& x)
{
    const S r = sqrt_(x.v);
    return Dual<S>(r, x.d / (S(2) * r));
}

EFEM_HD double valueOf(double x) { return x; }
template <typename S> EFEM_HD S valueOf(const Dual<S>& x) { return x.v; }

// log|J| from J - 1, accurate for J near 1.  The reference's invariants take
// (J^2)^(1/3) (I1bar = I1 / I3^(1/3), I3 = J^2: NeoHookean.cpp:9-10), which
// is |J|^(2/3) and stays finite for an inverted element (J < 0):
// log(-J) = log1p(-2 - (J - 1)).
template <typename D>
EFEM_HD D logAbsJ(const D& jm1)
{
    return valueOf(jm1) > -1.0 ? log1p_(jm1) : log1p_(-2.0 - jm1);
}

// J - 1 from H without cancellation (exact identity det(I+H) - 1)
template <typename D>
EFEM_HD D detIpHm1(const D* h)
{
    const D tr = h[0] + h[4] + h[8];
    const D trH2 = h[0] * h[0] + h[4] * h[4] + h[8] * h[8] +
                   2.0 * (h[1] * h[3] + h[2] * h[6] + h[5] * h[7]);
    const D det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                  h[2] * (h[3] * h[7] - h[4] * h[6]);
    return tr + 0.5 * (tr * tr - trH2) + det;
}

// P(F = I + H) for law with p = {C10, C01|C20, C30, D1}.  D is the type of
// the FP64-critical quantities (H, J - 1, J^-2/3 - 1, the volumetric
// coefficient): double, or Dual<double> for tangent products.  T is the type
// of the deviatoric terms and the cofactor: double or float (or their duals).
template <typename D, typename T>
EFEM_HD void piolaGen(int law, const double* p, const D* hd, T* P)
{
    T H[9];
    for (int k = 0; k < 9; ++k) H[k] = static_cast<T>(hd[k]);
    const D jm1 = detIpHm1(hd);
    const D J = 1.0 + jm1;
    // |J|^-2/3 - 1 = expm1(-2/3 log|J|), accurate for J near 1, consumed in
    // T.  With FP32 stresses and the Neo-Hookean or Mooney-Rivlin law it only
    // enters J^-2/3 = 1 + (J^-2/3 - 1), so it is formed in FP32 from the FP64
    // J - 1 (the FP64-critical step is J - 1 itself, free of cancellation;
    // the FP32 error of the small difference stays below an ulp of J^-2/3).
    // Yeoh's t = 3 (J^-2/3 - 1) + tr E J^-2/3 cancels to second order, so
    // there (and with FP64 stresses) it is formed in D
    D em1 = D(0.0);  // formed in D where Yeoh's t needs it
    T jm23;
    if constexpr (std::is_same<D, double>::value && std::is_same<T, float>::value) {
        if (law != LAW_YEOH) {
            const float j1 = static_cast<float>(jm1);
            jm23 = 1.0f + expm1f(-2.0f / 3.0f * (j1 > -1.0f ? log1pf(j1) : log1pf(-2.0f - j1)));
        } else {
            em1 = expm1_(-2.0 / 3.0 * logAbsJ(jm1));
            jm23 = static_cast<T>(1.0 + em1);
        }
    } else {
        em1 = expm1_(-2.0 / 3.0 * logAbsJ(jm1));
        jm23 = static_cast<T>(1.0 + em1);
    }
    // 1 / J and the volumetric coefficient (2/D1)(J - 1) formed in D, then
    // rounded once: the volumetric term dominates near incompressibility
    const T invJ = static_cast<T>(1.0 / J);
    const T kv = static_cast<T>(2.0 / p[3] * jm1);
    T F[9];
    for (int k = 0; k < 9; ++k) F[k] = H[k];
    F[0] += T(1);
    F[4] += T(1);
    F[8] += T(1);
    // cofactor matrix cof(F) = J F^-T
    T cf[9];
    cf[0] = F[4] * F[8] - F[5] * F[7];
    cf[1] = F[5] * F[6] - F[3] * F[8];
    cf[2] = F[3] * F[7] - F[4] * F[6];
    cf[3] = F[2] * F[7] - F[1] * F[8];
    cf[4] = F[0] * F[8] - F[2] * F[6];
    cf[5] = F[1] * F[6] - F[0] * F[7];
    cf[6] = F[1] * F[5] - F[2] * F[4];
    cf[7] = F[2] * F[3] - F[0] * F[5];
    cf[8] = F[0] * F[4] - F[1] * F[3];
    // E = H + H^T + H^T H (symmetric)
    T Em[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Em[3 * i + j] = H[3 * i + j] + H[3 * j + i] + H[i] * H[j] + H[3 + i] * H[3 + j] + H[6 + i] * H[6 + j];
    const T trE = Em[0] + Em[4] + Em[8];
    T A[9];  // deviatoric argument of the I1bar (and I2bar) terms
    for (int k = 0; k < 9; ++k) A[k] = Em[k];
    T c1 = static_cast<T>(p[0]);
    if (law == LAW_YEOH) {
        // t = I1 J^-2/3 - 3 = 3 (J^-2/3 - 1) + tr E J^-2/3
        const T t = static_cast<T>(3.0 * em1) + trE * jm23;
        c1 = static_cast<T>(p[0]) + T(2) * static_cast<T>(p[1]) * t + T(3) * static_cast<T>(p[2]) * t * t;
    }
    T s1 = T(2) * c1 * jm23 * invJ;  // 2 c1 J^-2/3 / J  (F^-T = cof / J)
    T dev[9];
    {
        const T m = trE * T(1.0 / 3.0);
        for (int k = 0; k < 9; ++k) dev[k] = s1 * A[k];
        dev[0] -= s1 * m;
        dev[4] -= s1 * m;
        dev[8] -= s1 * m;
    }
    if (law == LAW_MR) {
        // + 2 C01 J^-4/3 / J dev((1 + tr E) E - E^2)
        T E2[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                E2[3 * i + j] = Em[3 * i] * Em[j] + Em[3 * i + 1] * Em[3 + j] + Em[3 * i + 2] * Em[6 + j];
        T B[9];
        for (int k = 0; k < 9; ++k) B[k] = (T(1) + trE) * Em[k] - E2[k];
        const T mB = (B[0] + B[4] + B[8]) * T(1.0 / 3.0);
        const T s2 = T(2) * static_cast<T>(p[1]) * jm23 * jm23 * invJ;
        for (int k = 0; k < 9; ++k) dev[k] += s2 * B[k];
        dev[0] -= s2 * mB;
        dev[4] -= s2 * mB;
        dev[8] -= s2 * mB;
    }
    // P = cof(F) dev + (2/D1)(J-1) cof(F)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            P[3 * i + j] = cf[3 * i] * dev[j] + cf[3 * i + 1] * dev[3 + j] + cf[3 * i + 2] * dev[6 + j] +
                           kv * cf[3 * i + j];
}

template <typename T>
EFEM_HD void piolaFromHd(int law, const double* p, const double* hd, T* P)
{
    piolaGen<double, T>(law, p, hd, P);
}

template <typename T>
EFEM_HD void piolaFromH(int law, const double* p, const T* H, T* P)
{
    double hd[9];
    for (int k = 0; k < 9; ++k) hd[k] = static_cast<double>(H[k]);
    piolaFromHd<T>(law, p, hd, P);
}

// tangent product dP = (dP/dF)[dH] at F = I + H, FP64-critical part in FP64
template <typename T>
EFEM_HD void piolaTangent(int law, const double* p, const double* hd, const double* dhd, T* dP)
{
    Dual<double> h[9];
    for (int k = 0; k < 9; ++k) h[k] = Dual<double>(hd[k], dhd[k]);
    Dual<T> P[9];
    piolaGen<Dual<double>, Dual<T>>(law, p, h, P);
    for (int k = 0; k < 9; ++k) dP[k] = P[k].d;
}

// Derivatives of P with respect to the four material parameters, in the
// same slots as p.  P is linear in C10, C01 (MR), C20, C30 (Yeoh) -- the
// derivative is P with that parameter 1 and the others (and 1/D1) 0 -- and
// in 1/D1 through the volumetric term: dP/dD1 = -P_vol / D1.
template <typename T>
EFEM_HD void piolaParamDerivative(int law, const double* p, const double* hd, int k, T* dP)
{
    const double inf = HUGE_VAL;
    if (k == 3) {
        const double q[4] = {0.0, 0.0, 0.0, p[3]};
        piolaGen<double, T>(law, q, hd, dP);
        const T s = static_cast<T>(-1.0 / p[3]);
        for (int i = 0; i < 9; ++i) dP[i] = s * dP[i];
        return;
    }
    double q[4] = {0.0, 0.0, 0.0, inf};
    q[k] = 1.0;
    if ((law == LAW_NH && k != 0) || (law == LAW_MR && k == 2)) {
        for (int i = 0; i < 9; ++i) dP[i] = T(0);
        return;
    }
    piolaGen<double, T>(law, q, hd, dP);
}

// strain energy density at F = I + H (FP64, or its dual)
template <typename D>
EFEM_HD D energyGen(int law, const double* p, const D* H)
{
    const D jm1 = detIpHm1(H);
    const D em1 = expm1_(-2.0 / 3.0 * logAbsJ(jm1));
    D E[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            E[3 * i + j] = H[3 * i + j] + H[3 * j + i] + H[i] * H[j] + H[3 + i] * H[3 + j] + H[6 + i] * H[6 + j];
    const D trE = E[0] + E[4] + E[8];
    const D t = 3.0 * em1 + trE * (1.0 + em1);  // I1bar - 3
    D W = jm1 * jm1 / p[3];
    if (law == LAW_YEOH) return W + p[0] * t + p[1] * t * t + p[2] * t * t * t;
    W += p[0] * t;
    if (law == LAW_MR) {
        // I2 - 3 = 2 tr E + (tr^2 E - tr E^2)/2 ; I2bar - 3 = (I2 - 3) J^-4/3 + 3 (J^-4/3 - 1)
        D trE2 = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) trE2 += E[3 * i + j] * E[3 * j + i];
        const D i2m3 = 2.0 * trE + 0.5 * (trE * trE - trE2);
        const D j43 = (1.0 + em1) * (1.0 + em1);
        W += p[1] * (i2m3 * j43 + 3.0 * (j43 - 1.0));
    }
    return W;
}

EFEM_HD double energyFromH(int law, const double* p, const double* H) { return energyGen<double>(law, p, H); }

}  // namespace efem
