// phc_solver_utils.hpp -- SolverUtils (src/utils/utils.hpp:9-59): the Ricker
// wavelet and the source-term samples, on libphc.so's setup routines.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "phc.h"
#include "phc_setup.h"

struct SolverUtils {
    // The reference's FD translation unit binds ::exp(double) in
    // evaluateRicker, its SEM one std::exp(float) (utils.hpp:31-44 under each
    // unit's using-directives): PHC_EXP_DOUBLE for FD, PHC_EXP_FLOAT for SEM.
    int expVariant = PHC_EXP_FLOAT;

    // utils.hpp:11-46
    float evaluateRicker(float const &time_n, float const &f0, int order)
    {
        return phc_ricker(time_n, f0, order, expVariant);
    }
    // utils.hpp:48-57
    std::vector<float> computeSourceTerm(const int nSamples, const float timeSample, const float f0, const int order)
    {
        std::vector<float> st((size_t)(nSamples > 0 ? nSamples : 0));
        if (phc_source_term(nSamples, timeSample, f0, order, expVariant, st.data()) != PHC_OK)
            throw std::runtime_error(std::string("computeSourceTerm: ") + phc_last_error());
        return st;
    }
};
