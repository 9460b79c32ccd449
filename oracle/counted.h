// Op-counting scalars for the oracle's counting build -- TEST INFRASTRUCTURE.
//
// The Makefile's liboracle_count.so compiles the oracle's sources with
// `-include counted.h`: after the standard headers below, `float` and
// `double` name Counted<float> and Counted<double>, which compute exactly as
// the built-in types and tally every arithmetic operation the oracle
// evaluates while counting is on (cnt::on):
//   +, -, *, /, comparisons, min / max, sqrt   1 flop each (cnt::flops)
//   acos, cos, sin, atan2, pow, cbrt, exp, log  1 SFU op each (cnt::sfu)
// Negation, fabs and conversions are free.  Counts go to the bucket
// cnt::phase (the caller's split into forward, interior and boundary work).
// Counted<R> holds one R and nothing else, so arrays and structs of them
// have the built-in layout and the C ABI of liboracle.so is unchanged.
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <stddef.h>
#include <stdint.h>
#include <math.h>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace cnt {
constexpr int PHASES = 4;
extern bool on;
extern int phase;
extern uint64_t flops[PHASES], sfu[PHASES];
inline void op(uint64_t n = 1) {
    if (on) flops[phase] += n;
}
inline void sf(uint64_t n = 1) {
    if (on) sfu[phase] += n;
}
}  // namespace cnt

template <class R>
struct Counted {
    R v;
    Counted() = default;
    template <class T, std::enable_if_t<std::is_arithmetic_v<T>, int> = 0>
    constexpr Counted(T x) : v((R)x) {}
    template <class S>
    constexpr Counted(Counted<S> x) : v((R)x.v) {}
    template <class T, std::enable_if_t<std::is_arithmetic_v<T>, int> = 0>
    explicit constexpr operator T() const { return (T)v; }
    Counted operator-() const { return Counted(-v); }
    Counted operator+() const { return *this; }
    template <class B>
    Counted &operator+=(B b);
    template <class B>
    Counted &operator-=(B b);
    template <class B>
    Counted &operator*=(B b);
    template <class B>
    Counted &operator/=(B b);
};

namespace cnt {
template <class T>
struct is_c : std::false_type {};
template <class R>
struct is_c<Counted<R>> : std::true_type {};
template <class T>
constexpr bool ok = is_c<T>::value || std::is_arithmetic_v<T>;
template <class A, class B>
constexpr bool mixed = (is_c<A>::value || is_c<B>::value) && ok<A> && ok<B>;
template <class T>
constexpr auto raw(T x) {
    if constexpr (is_c<T>::value) return x.v;
    else return x;
}
template <class A, class B>
using res = Counted<decltype(raw(std::declval<A>()) + raw(std::declval<B>()))>;
}  // namespace cnt

#define CNT_BINOP(OP)                                                     \
    template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0> \
    inline cnt::res<A, B> operator OP(A a, B b) {                         \
        cnt::op();                                                        \
        return cnt::res<A, B>(cnt::raw(a) OP cnt::raw(b));                \
    }
CNT_BINOP(+)
CNT_BINOP(-)
CNT_BINOP(*)
CNT_BINOP(/)
#undef CNT_BINOP
#define CNT_CMP(OP)                                                       \
    template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0> \
    inline bool operator OP(A a, B b) {                                   \
        cnt::op();                                                        \
        return cnt::raw(a) OP cnt::raw(b);                                \
    }
CNT_CMP(<)
CNT_CMP(>)
CNT_CMP(<=)
CNT_CMP(>=)
CNT_CMP(==)
CNT_CMP(!=)
#undef CNT_CMP

template <class R>
template <class B>
Counted<R> &Counted<R>::operator+=(B b) { return *this = Counted<R>(*this + b); }
template <class R>
template <class B>
Counted<R> &Counted<R>::operator-=(B b) { return *this = Counted<R>(*this - b); }
template <class R>
template <class B>
Counted<R> &Counted<R>::operator*=(B b) { return *this = Counted<R>(*this * b); }
template <class R>
template <class B>
Counted<R> &Counted<R>::operator/=(B b) { return *this = Counted<R>(*this / b); }

namespace std {
template <class R>
struct numeric_limits<Counted<R>> : numeric_limits<R> {
    static constexpr Counted<R> infinity() { return numeric_limits<R>::infinity(); }
    static constexpr Counted<R> max() { return numeric_limits<R>::max(); }
    static constexpr Counted<R> min() { return numeric_limits<R>::min(); }
    static constexpr Counted<R> lowest() { return numeric_limits<R>::lowest(); }
    static constexpr Counted<R> epsilon() { return numeric_limits<R>::epsilon(); }
    static constexpr Counted<R> quiet_NaN() { return numeric_limits<R>::quiet_NaN(); }
};
#define CNT_F1(NAME, KIND)                           \
    template <class R>                               \
    inline Counted<R> NAME(Counted<R> x) {           \
        KIND;                                        \
        return Counted<R>(std::NAME(x.v));           \
    }
CNT_F1(sqrt, cnt::op())
CNT_F1(fabs, (void)0)
CNT_F1(abs, (void)0)
CNT_F1(floor, (void)0)
CNT_F1(ceil, (void)0)
CNT_F1(trunc, (void)0)
CNT_F1(round, (void)0)
CNT_F1(acos, cnt::sf())
CNT_F1(asin, cnt::sf())
CNT_F1(atan, cnt::sf())
CNT_F1(cos, cnt::sf())
CNT_F1(sin, cnt::sf())
CNT_F1(exp, cnt::sf())
CNT_F1(log, cnt::sf())
CNT_F1(cbrt, cnt::sf())
#undef CNT_F1
template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0>
inline cnt::res<A, B> pow(A a, B b) {
    cnt::sf();
    return cnt::res<A, B>(std::pow(cnt::raw(a), cnt::raw(b)));
}
template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0>
inline cnt::res<A, B> atan2(A a, B b) {
    cnt::sf();
    return cnt::res<A, B>(std::atan2(cnt::raw(a), cnt::raw(b)));
}
template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0>
inline cnt::res<A, B> fmin(A a, B b) {
    cnt::op();
    return cnt::res<A, B>(std::fmin(cnt::raw(a), cnt::raw(b)));
}
template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0>
inline cnt::res<A, B> fmax(A a, B b) {
    cnt::op();
    return cnt::res<A, B>(std::fmax(cnt::raw(a), cnt::raw(b)));
}
template <class A, class B, std::enable_if_t<cnt::mixed<A, B>, int> = 0>
inline cnt::res<A, B> fma(A a, B b, cnt::res<A, B> c) {
    cnt::op(2);
    return cnt::res<A, B>(std::fma(cnt::raw(a), cnt::raw(b), c.v));
}
template <class R>
inline bool isfinite(Counted<R> x) { return std::isfinite(x.v); }
template <class R>
inline bool isnan(Counted<R> x) { return std::isnan(x.v); }
template <class R>
inline bool isinf(Counted<R> x) { return std::isinf(x.v); }
template <class R>
inline bool signbit(Counted<R> x) { return std::signbit(x.v); }
template <class R>
inline Counted<R> copysign(Counted<R> a, Counted<R> b) { return Counted<R>(std::copysign(a.v, b.v)); }
}  // namespace std
using std::sqrt;
using std::fabs;

#define float Counted<float>
#define double Counted<double>
