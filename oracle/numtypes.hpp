// Number types for the CPU oracle.  TEST INFRASTRUCTURE ONLY: nothing in the
// shipped library links this; it is the checker for the HIP kernels.
//
// The restatement is written once as templates over a real type R:
//   float            -> the FP32 restatement (reference operation order,
//                       -ffp-contract=off so every + - * / rounds separately)
//   double           -> the FP64 accuracy oracle
//   Counted<float>   -> the op-counting build: every + - * / sqrt exp acos
//                       asin erf counts 1 FLOP; comparisons, fabs, min/max,
//                       conversions count 0; trig of fixed rotation angles
//                       (rot_cos / rot_sin) counts 0.
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

struct OpCounter {
    static thread_local uint64_t flops;
    static thread_local bool enabled;
};

template <class T>
struct Counted {
    T v;
    Counted() : v(0) {}
    Counted(T x) : v(x) {}  // NOLINT: implicit like a builtin real
    template <class U>
    explicit Counted(Counted<U> o) : v(static_cast<T>(o.v)) {}
    explicit operator T() const { return v; }
    static void tick(uint64_t n = 1) {
        if (OpCounter::enabled) OpCounter::flops += n;
    }
};

#define ORACLE_CNT_BIN(op)                                                         \
    template <class T>                                                             \
    inline Counted<T> operator op(Counted<T> a, Counted<T> b) {                    \
        Counted<T>::tick();                                                        \
        return Counted<T>(a.v op b.v);                                             \
    }                                                                              \
    template <class T>                                                             \
    inline Counted<T>& operator op##=(Counted<T>& a, Counted<T> b) {               \
        Counted<T>::tick();                                                        \
        a.v = a.v op b.v;                                                          \
        return a;                                                                  \
    }
ORACLE_CNT_BIN(+)
ORACLE_CNT_BIN(-)
ORACLE_CNT_BIN(*)
ORACLE_CNT_BIN(/)
#undef ORACLE_CNT_BIN

template <class T> inline Counted<T> operator-(Counted<T> a) { return Counted<T>(-a.v); }
#define ORACLE_CNT_CMP(op) \
    template <class T> inline bool operator op(Counted<T> a, Counted<T> b) { return a.v op b.v; }
ORACLE_CNT_CMP(<)
ORACLE_CNT_CMP(>)
ORACLE_CNT_CMP(<=)
ORACLE_CNT_CMP(>=)
ORACLE_CNT_CMP(==)
ORACLE_CNT_CMP(!=)
#undef ORACLE_CNT_CMP

// ---- math, overloaded so the templates read like plain C math ------------
inline float r_sqrt(float x) { return std::sqrt(x); }
inline double r_sqrt(double x) { return std::sqrt(x); }
inline float r_exp(float x) { return std::exp(x); }
inline double r_exp(double x) { return std::exp(x); }
inline float r_fabs(float x) { return std::fabs(x); }
inline double r_fabs(double x) { return std::fabs(x); }
inline float r_acos(float x) { return std::acos(x); }
inline double r_acos(double x) { return std::acos(x); }
inline float r_asin(float x) { return std::asin(x); }
inline double r_asin(double x) { return std::asin(x); }
inline float r_erf(float x) { return std::erf(x); }
inline double r_erf(double x) { return std::erf(x); }
inline float r_cos(float x) { return std::cos(x); }
inline double r_cos(double x) { return std::cos(x); }
inline float r_sin(float x) { return std::sin(x); }
inline double r_sin(double x) { return std::sin(x); }
inline float rot_cos(float x) { return std::cos(x); }
inline double rot_cos(double x) { return std::cos(x); }
inline float rot_sin(float x) { return std::sin(x); }
inline double rot_sin(double x) { return std::sin(x); }

#define ORACLE_CNT_FN(name, cost)                                   \
    template <class T> inline Counted<T> name(Counted<T> a) {       \
        Counted<T>::tick(cost);                                     \
        return Counted<T>(name(a.v));                               \
    }
ORACLE_CNT_FN(r_sqrt, 1)
ORACLE_CNT_FN(r_exp, 1)
ORACLE_CNT_FN(r_fabs, 0)
ORACLE_CNT_FN(r_acos, 1)
ORACLE_CNT_FN(r_asin, 1)
ORACLE_CNT_FN(r_erf, 1)
ORACLE_CNT_FN(r_cos, 1)
ORACLE_CNT_FN(r_sin, 1)
ORACLE_CNT_FN(rot_cos, 0)
ORACLE_CNT_FN(rot_sin, 0)
#undef ORACLE_CNT_FN

// Wide type used where the reference mixes in a double literal (M_PI, sqrt(2.0)):
// float -> double, double -> double, Counted<float> -> Counted<double>.
template <class R> struct Wide { using type = double; };
template <class T> struct Wide<Counted<T>> { using type = Counted<double>; };

template <class R> inline double to_double(R x) { return static_cast<double>(x); }
template <class T> inline double to_double(Counted<T> x) { return static_cast<double>(x.v); }

}  // namespace oracle
