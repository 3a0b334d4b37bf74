// Hyperelastic PK1 stress for the three laws of src/material/*.cpp, written
// in terms of the displacement gradient H = F - I so that the FP32 path does
// not cancel:
//   * J - 1 = tr H + (tr^2 H - tr H^2)/2 + det H, always in FP64;
//   * J F^-T = cof F (no division by J for the volumetric term);
//   * P = c1 J^-2/3 (2F - 2/3 I1 cof F / J)                     [I1bar term]
//       + C01 J^-4/3 (2F(I1 I - C) - 4/3 I2 cof F / J)          [I2bar term, MR]
//       + (2/D1)(J - 1) cof F                                   [volumetric]
// which equals NeoHookean.cpp:4-20, MooneyRivlin.cpp:4-24 and Yeoh.cpp:4-20
// (dJ/dF = J F^-T, I1bar = I1 / I3^(1/3), I2bar = I2 / I3^(2/3)).
// Strain energies (no counterpart in the reference, which never evaluates
// one): W = c(I1bar-3)[+C01(I2bar-3)] + (J-1)^2/D1, Yeoh sum_i Ci0(I1bar-3)^i.
#pragma once

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
    if (law == LAW_NH) return 2.94001 * p[0] + 2.0 / p[3];
    if (law == LAW_MR) return 3.818 * p[1] + 2.94001 * p[0] + 2.0 / p[3];
    return 2.94001 * p[0] + 3.51194 * p[1] + 3.3154 * p[2] + 2.0 / p[3];
}

// J - 1 from H without cancellation (exact identity det(I+H) - 1)
EFEM_HD double detIpHm1(const double* h)
{
    const double tr = h[0] + h[4] + h[8];
    const double trH2 = h[0] * h[0] + h[4] * h[4] + h[8] * h[8] +
                        2.0 * (h[1] * h[3] + h[2] * h[6] + h[5] * h[7]);
    const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                       h[2] * (h[3] * h[7] - h[4] * h[6]);
    return tr + 0.5 * (tr * tr - trH2) + det;
}

// P(F = I + H) for law with p = {C10, C01|C20, C30, D1}.  T = double or float;
// J - 1 and the volumetric coefficient are FP64 in both cases.
template <typename T>
EFEM_HD void piolaFromH(int law, const double* p, const T* H, T* P)
{
    double hd[9];
    for (int k = 0; k < 9; ++k) hd[k] = static_cast<double>(H[k]);
    const double jm1 = detIpHm1(hd);
    const double J = 1.0 + jm1;
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
    T I1 = T(0);
    for (int k = 0; k < 9; ++k) I1 += F[k] * F[k];
    const T Jt = static_cast<T>(J);
    const T invJ = T(1) / Jt;
    const T jm23 = T(1) / static_cast<T>(std::cbrt(J * J));  // J^-2/3
    T c1 = static_cast<T>(p[0]);
    if (law == LAW_YEOH) {
        const T t = I1 * jm23 - T(3);
        c1 = static_cast<T>(p[0]) + T(2) * static_cast<T>(p[1]) * t +
             T(3) * static_cast<T>(p[2]) * t * t;
    }
    const T a = c1 * jm23;
    const T b = T(2.0 / 3.0) * I1 * invJ;
    const double kvol = 2.0 / p[3] * jm1;  // (2/D1)(J-1)
    const T kv = static_cast<T>(kvol);
    for (int k = 0; k < 9; ++k) P[k] = a * (T(2) * F[k] - b * cf[k]) + kv * cf[k];
    if (law == LAW_MR) {
        // C = F^T F ; I2 = (I1^2 - tr C^2)/2
        T C[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
        T trCC = T(0);
        for (int k = 0; k < 9; ++k) trCC += C[k] * C[k];
        const T I2 = T(0.5) * (I1 * I1 - trCC);
        const T jm43 = jm23 * jm23;
        const T c01 = static_cast<T>(p[1]) * jm43;
        const T b2 = T(4.0 / 3.0) * I2 * invJ;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                // (F (I1 I - C))_ij
                const T fm = I1 * F[3 * i + j] -
                             (F[3 * i] * C[j] + F[3 * i + 1] * C[3 + j] + F[3 * i + 2] * C[6 + j]);
                P[3 * i + j] += c01 * (T(2) * fm - b2 * cf[3 * i + j]);
            }
    }
}

// strain energy density at F = I + H (FP64)
EFEM_HD double energyFromH(int law, const double* p, const double* H)
{
    const double jm1 = detIpHm1(H);
    const double J = 1.0 + jm1;
    double F[9];
    for (int k = 0; k < 9; ++k) F[k] = H[k];
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    double I1 = 0.0;
    for (int k = 0; k < 9; ++k) I1 += F[k] * F[k];
    const double jm23 = 1.0 / std::cbrt(J * J);
    const double i1b3 = I1 * jm23 - 3.0;
    double W = jm1 * jm1 / p[3];
    if (law == LAW_YEOH) {
        W += p[0] * i1b3 + p[1] * i1b3 * i1b3 + p[2] * i1b3 * i1b3 * i1b3;
    } else {
        W += p[0] * i1b3;
        if (law == LAW_MR) {
            double C[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
            double trCC = 0.0;
            for (int k = 0; k < 9; ++k) trCC += C[k] * C[k];
            const double I2 = 0.5 * (I1 * I1 - trCC);
            W += p[1] * (I2 * jm23 * jm23 - 3.0);
        }
    }
    return W;
}

}  // namespace efem
