// The reuse policy of the C++ backend against the cases of the reference's policy test
// (test/test_reuse.cpp:50-110): iterations per decade, the break-even rule, the safety nets, the
// lower-median costs, and the Off and Refactor modes. CPU only: no GPU call.
#include "s2l_backend.hpp"

#include <cmath>
#include <cstdio>
#include <stdexcept>

using s2l::prec::ReusePolicy;
using s2l::prec::UpdateLevel;

static int failures = 0;
static void check(bool ok, const char* what)
{
    std::printf("%s: %s\n", ok ? "PASS" : "FAIL", what);
    if (!ok) ++failures;
}
static bool near(double a, double b) { return std::abs(a - b) < 1e-12; }

int main()
{
    using P = ReusePolicy;
    check(near(P::perDecade(10, 1e-5), 2.0) && P::perDecade(0, 1e-5) == 0.0 && P::perDecade(5, 1.0) == 0.0,
          "iterations per decade");

    // rebuild 2.0 s, refactorize 0.2 s: premium 1.8 s; q_ref = 10 / 8 = 1.25, 0.1 s per iteration
    P a(P::Mode::Auto, 0.0, 0);
    a.record(UpdateLevel::Rebuild, 10, 1e-8, true, 2.0, 1.0);
    bool ok = a.decide(false) == UpdateLevel::Reuse && a.decide(true) == UpdateLevel::Refactorize && a.wasted() == 0.0;
    a.record(UpdateLevel::Refactorize, 15, 1e-8, true, 0.2, 1.5);
    ok = ok && near(a.wasted(), 0.5) && a.decide(true) == UpdateLevel::Refactorize;
    a.record(UpdateLevel::Reuse, 15, 1e-8, true, 0.0, 1.5);
    a.record(UpdateLevel::Refactorize, 19, 1e-8, true, 0.2, 1.9);
    ok = ok && near(a.wasted(), 1.9) && near(a.rebuildPremium(), 1.8);
    ok = ok && a.decide(true) == UpdateLevel::Rebuild && a.decide(false) == UpdateLevel::Reuse;
    a.record(UpdateLevel::Rebuild, 12, 1e-8, true, 2.5, 1.2);
    ok = ok && a.decide(true) == UpdateLevel::Refactorize && a.wasted() == 0.0 && near(a.reference(), 1.5);
    check(ok, "auto: break-even rule");

    P g(P::Mode::Auto, 1.5, 0);
    g.record(UpdateLevel::Rebuild, 10, 1e-8, true, 1e9, 1.0);
    g.record(UpdateLevel::Refactorize, 16, 1e-8, true, 0.1, 1.6);
    ok = g.decide(true) == UpdateLevel::Rebuild;
    P n(P::Mode::Auto, 0.0, 0);
    n.record(UpdateLevel::Rebuild, 10, 1e-8, true, 1e9, 1.0);
    n.record(UpdateLevel::Refactorize, 500, 1e-3, false, 0.1, 50.0);
    ok = ok && n.decide(true) == UpdateLevel::Rebuild;
    P b(P::Mode::Auto, 0.0, 2);
    b.record(UpdateLevel::Rebuild, 10, 1e-8, true, 1e9, 1.0);
    ok = ok && b.decide(true) == UpdateLevel::Refactorize;
    b.record(UpdateLevel::Refactorize, 10, 1e-8, true, 0.1, 1.0);
    ok = ok && b.decide(true) == UpdateLevel::Refactorize;
    b.record(UpdateLevel::Refactorize, 10, 1e-8, true, 0.1, 1.0);
    ok = ok && b.decide(true) == UpdateLevel::Rebuild;
    check(ok, "auto: safety nets (growth, age, no convergence)");

    ok = P::lowerMedian({}) == 0.0 && P::lowerMedian({3.0}) == 3.0 && P::lowerMedian({5.0, 1.0}) == 1.0 &&
         P::lowerMedian({4.0, 1.0, 2.0}) == 2.0 && P::lowerMedian({4.0, 1.0, 2.0, 3.0}) == 2.0;
    P s(P::Mode::Auto, 0.0, 0);
    s.record(UpdateLevel::Rebuild, 10, 1e-8, true, 2.0, 1.0);
    s.record(UpdateLevel::Refactorize, 10, 1e-8, true, 0.2, 1.0);
    s.record(UpdateLevel::Rebuild, 10, 1e-8, true, 30.0, 1.0);
    ok = ok && near(s.rebuildPremium(), 1.8);
    s.record(UpdateLevel::Refactorize, 10, 1e-8, true, 9.0, 1.0);
    ok = ok && near(s.rebuildPremium(), 1.8);
    check(ok, "auto: lower-median costs ignore one stall");

    P off(P::Mode::Off, 0.0, 0);
    check(off.decide(false) == UpdateLevel::Rebuild && off.decide(true) == UpdateLevel::Rebuild, "off: rebuild before every solve");
    P r(P::Mode::Refactor, 0.0, 0);
    r.record(UpdateLevel::Rebuild, 10, 1e-8, true, 2.0, 1.0);
    r.record(UpdateLevel::Refactorize, 900, 1e-3, false, 0.1, 90.0);
    check(r.decide(false) == UpdateLevel::Reuse && r.decide(true) == UpdateLevel::Refactorize, "refactor: never rebuild");
    bool threw = false;
    try {
        (void)P::parseMode("sometimes");
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw && P::parseMode("auto") == P::Mode::Auto, "mode names");
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
