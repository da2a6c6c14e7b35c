// ORACLE — test infrastructure only. FLOP-counting scalar for a second build
// of the oracle (liboracle_flops.so, `make -C oracle flops`): every source is
// compiled with this header force-included, which ends by defining `double`
// as FlopReal. FlopReal holds one IEEE double and has the layout and calling
// convention of a double (one trivially copyable member), so the C-ABI of the
// library is unchanged; each arithmetic operation it performs is counted.
//
// Counting rules (the algorithmic FLOP convention of the roofline figure):
// + - * / and sqrt count 1 each (the oracle is built with -ffp-contract=off,
// so a*b + c is a multiply and an add: 2, as an FMA counts 2); fma counts 2;
// sin, cos, exp, log, pow, atan2, asin, acos, atan, hypot are counted apart as
// transcendental calls; comparisons, fabs, fmin, fmax, floor, fmod, copysign
// and conversions are not counted. Counts go to the category in force
// (ORACLE_FLOP_SCOPE, flop_scope.h): the per-object force model and step, the epoch-only
// inputs, or the NRLMSISE-00 density.
#pragma once
#define ORACLE_FLOP_COUNT 1

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <initializer_list>
#include <iostream>
#include <istream>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>
#include <omp.h>

typedef __typeof__(1.0) f64;   // the builtin double (the macro below renames the keyword)

namespace oracle_flop {
enum Category { kPerObject = 0, kEpoch = 1, kNrlmsise = 2, kCategories = 3 };
struct Counts {
    unsigned long long flops[kCategories];
    unsigned long long trans[kCategories];
    unsigned long long calls[kCategories];   // entries into a scope of the category
};
extern thread_local Counts counts;
extern thread_local int category;
inline void add(unsigned long long n) { counts.flops[category] += n; }
inline void addTrans() { counts.trans[category] += 1; }
// Scope guard: counts inside go to category c (the innermost scope wins).
struct Scope {
    int saved;
    explicit Scope(int c) : saved(category) {
        category = c;
        counts.calls[c] += 1;
    }
    ~Scope() { category = saved; }
};
}  // namespace oracle_flop

struct FlopReal {
    f64 v;
    FlopReal() = default;
    constexpr FlopReal(f64 x) : v(x) {}
    constexpr FlopReal(float x) : v(x) {}
    constexpr FlopReal(int x) : v(x) {}
    constexpr FlopReal(long x) : v(f64(x)) {}
    constexpr FlopReal(long long x) : v(f64(x)) {}
    constexpr FlopReal(unsigned x) : v(x) {}
    constexpr FlopReal(unsigned long x) : v(f64(x)) {}
    constexpr FlopReal(unsigned long long x) : v(f64(x)) {}
    explicit constexpr operator f64() const { return v; }
    explicit constexpr operator float() const { return float(v); }
    explicit constexpr operator int() const { return int(v); }
    explicit constexpr operator long() const { return long(v); }
    explicit constexpr operator long long() const { return (long long)v; }
    explicit constexpr operator unsigned() const { return unsigned(v); }
    explicit constexpr operator unsigned long() const { return (unsigned long)v; }
    explicit constexpr operator unsigned long long() const { return (unsigned long long)v; }
    explicit constexpr operator bool() const { return v != 0.0; }
    FlopReal& operator+=(FlopReal o) { oracle_flop::add(1); v += o.v; return *this; }
    FlopReal& operator-=(FlopReal o) { oracle_flop::add(1); v -= o.v; return *this; }
    FlopReal& operator*=(FlopReal o) { oracle_flop::add(1); v *= o.v; return *this; }
    FlopReal& operator/=(FlopReal o) { oracle_flop::add(1); v /= o.v; return *this; }
};
static_assert(sizeof(FlopReal) == sizeof(f64) && alignof(FlopReal) == alignof(f64), "FlopReal must be laid out as a double");

constexpr FlopReal operator+(FlopReal a, FlopReal b) {
    if (!__builtin_is_constant_evaluated()) oracle_flop::add(1);
    return FlopReal(a.v + b.v);
}
constexpr FlopReal operator-(FlopReal a, FlopReal b) {
    if (!__builtin_is_constant_evaluated()) oracle_flop::add(1);
    return FlopReal(a.v - b.v);
}
constexpr FlopReal operator*(FlopReal a, FlopReal b) {
    if (!__builtin_is_constant_evaluated()) oracle_flop::add(1);
    return FlopReal(a.v * b.v);
}
constexpr FlopReal operator/(FlopReal a, FlopReal b) {
    if (!__builtin_is_constant_evaluated()) oracle_flop::add(1);
    return FlopReal(a.v / b.v);
}
constexpr FlopReal operator-(FlopReal a) { return FlopReal(-a.v); }
constexpr FlopReal operator+(FlopReal a) { return a; }
constexpr bool operator<(FlopReal a, FlopReal b) { return a.v < b.v; }
constexpr bool operator>(FlopReal a, FlopReal b) { return a.v > b.v; }
constexpr bool operator<=(FlopReal a, FlopReal b) { return a.v <= b.v; }
constexpr bool operator>=(FlopReal a, FlopReal b) { return a.v >= b.v; }
constexpr bool operator==(FlopReal a, FlopReal b) { return a.v == b.v; }
constexpr bool operator!=(FlopReal a, FlopReal b) { return a.v != b.v; }
inline bool operator!(FlopReal a) { return a.v == 0.0; }
inline std::istream& operator>>(std::istream& s, FlopReal& x) { return s >> x.v; }
inline std::ostream& operator<<(std::ostream& s, FlopReal x) { return s << x.v; }

#define ORACLE_FLOP_UNARY(name, counted)                                  \
    inline FlopReal name(FlopReal x) {                                    \
        counted;                                                          \
        return FlopReal(::std::name(x.v));                                \
    }
#define ORACLE_FLOP_BINARY(name, counted)                                 \
    inline FlopReal name(FlopReal x, FlopReal y) {                        \
        counted;                                                          \
        return FlopReal(::std::name(x.v, y.v));                           \
    }
namespace oracle_flop_math {
ORACLE_FLOP_UNARY(sqrt, oracle_flop::add(1))
ORACLE_FLOP_UNARY(sin, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(cos, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(exp, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(log, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(log10, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(asin, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(acos, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(atan, oracle_flop::addTrans())
ORACLE_FLOP_UNARY(fabs, (void)0)
ORACLE_FLOP_UNARY(floor, (void)0)
ORACLE_FLOP_UNARY(ceil, (void)0)
ORACLE_FLOP_UNARY(round, (void)0)
ORACLE_FLOP_UNARY(trunc, (void)0)
ORACLE_FLOP_BINARY(atan2, oracle_flop::addTrans())
ORACLE_FLOP_BINARY(pow, oracle_flop::addTrans())
ORACLE_FLOP_BINARY(hypot, oracle_flop::addTrans())
ORACLE_FLOP_BINARY(fmod, (void)0)
ORACLE_FLOP_BINARY(fmin, (void)0)
ORACLE_FLOP_BINARY(fmax, (void)0)
ORACLE_FLOP_BINARY(copysign, (void)0)
}  // namespace oracle_flop_math

// The oracle calls the math functions as std::name; the overloads for
// FlopReal are made visible there.
namespace std {
using oracle_flop_math::sqrt;
using oracle_flop_math::sin;
using oracle_flop_math::cos;
using oracle_flop_math::exp;
using oracle_flop_math::log;
using oracle_flop_math::log10;
using oracle_flop_math::asin;
using oracle_flop_math::acos;
using oracle_flop_math::atan;
using oracle_flop_math::fabs;
using oracle_flop_math::floor;
using oracle_flop_math::ceil;
using oracle_flop_math::round;
using oracle_flop_math::trunc;
using oracle_flop_math::atan2;
using oracle_flop_math::pow;
using oracle_flop_math::hypot;
using oracle_flop_math::fmod;
using oracle_flop_math::fmin;
using oracle_flop_math::fmax;
using oracle_flop_math::copysign;
inline FlopReal abs(FlopReal x) { return FlopReal(::std::fabs(x.v)); }
inline FlopReal fma(FlopReal a, FlopReal b, FlopReal c) { oracle_flop::add(2); return FlopReal(::std::fma(a.v, b.v, c.v)); }
inline FlopReal ldexp(FlopReal x, int e) { return FlopReal(::std::ldexp(x.v, e)); }
inline bool isfinite(FlopReal x) { return ::std::isfinite(x.v); }
inline bool isnan(FlopReal x) { return ::std::isnan(x.v); }
inline bool isinf(FlopReal x) { return ::std::isinf(x.v); }
inline bool signbit(FlopReal x) { return ::std::signbit(x.v); }
inline const FlopReal& min(const FlopReal& a, const FlopReal& b) { return b < a ? b : a; }
inline const FlopReal& max(const FlopReal& a, const FlopReal& b) { return a < b ? b : a; }
inline string to_string(FlopReal x) { return to_string(x.v); }
template <>
class numeric_limits<FlopReal> : public numeric_limits<f64> {};
}  // namespace std

#define double FlopReal
