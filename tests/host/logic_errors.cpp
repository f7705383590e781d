// The host layer's configuration errors, which need no GPU: the reference's
// logic_error cases (MethodOfLinesTVD::setRungeKuttaOrder,
// src/SolutionSchemes.cpp:77-84) and the field count of NewtonianMHD with
// cooling (src/ConservationLaws.cpp:354-358,503-506).  Prints one line per
// check; exit status 0 when every check holds.
#include <cstdio>
#include <memory>
#include <stdexcept>

#include "../../mara2_amd/host/mara2.hpp"

using namespace mara2;

template <typename F> static int expect_logic_error(const char* what, F f)
{
    try
    {
        f();
    }
    catch (const std::logic_error&)
    {
        std::printf("ok   %s\n", what);
        return 0;
    }
    std::printf("FAIL %s: no logic_error\n", what);
    return 1;
}

template <typename F> static int expect_no_error(const char* what, F f)
{
    try
    {
        f();
    }
    catch (const std::exception& e)
    {
        std::printf("FAIL %s: %s\n", what, e.what());
        return 1;
    }
    std::printf("ok   %s\n", what);
    return 0;
}

int main()
{
    int bad = 0;
    auto ss = std::make_shared<GpuMethodOfLinesTVD>();
    bad += expect_logic_error("rk order 0", [&] { ss->setRungeKuttaOrder(0); });
    bad += expect_logic_error("rk order 4", [&] { ss->setRungeKuttaOrder(4); });
    bad += expect_no_error("rk orders 1..3", [&] {
        for (int k = 1; k <= 3; ++k) ss->setRungeKuttaOrder(k);
    });
    auto mhd = std::make_shared<NewtonianMHD>();
    bad += expect_no_error("MHD cooling adds the auxiliary field", [&] {
        mhd->setCoolingRate(0.1);
        if (mhd->getNumConserved() != 9 || mhd->getPrimitiveName(8) != "auxiliary1")
            throw std::runtime_error("expected 9 fields, the last auxiliary1");
        mhd->setCoolingRate(-1.0);
        if (mhd->getNumConserved() != 8) throw std::runtime_error("expected 8 fields with cooling off");
    });
    // a source function on a scheme whose law cools is refused before any
    // block is touched (the reference adds cooling to L before the
    // function's S, an order one S array cannot keep)
    auto cooled = std::make_shared<NewtonianMHD>();
    cooled->setCoolingRate(0.3);
    auto ss2 = std::make_shared<GpuMethodOfLinesTVD>();
    ss2->setFieldOperator(std::make_shared<FieldOperator>(cooled));
    bad += expect_logic_error("source with cooling", [&] {
        ss2->setSourceTermsCodeObject("any.hsaco", "any_kernel", {1.0});
    });
    bad += expect_no_error("removing a source with cooling", [&] { ss2->setSourceTermsCodeObject("", ""); });
    return bad == 0 ? 0 : 1;
}
