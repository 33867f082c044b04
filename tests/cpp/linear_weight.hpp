// A user exposure-weight function in the shape of enfuse's example plugin
// (doc/examples/enfuse/linear.cc): a triangle of half-width equal to the
// Gaussian's FWHM around the optimum, 1 at the optimum and 0 beyond.
#pragma once

#include <cmath>

#include "../../include/enblend_amd/exposure_weight.hpp"

struct LinearWeight : public ExposureWeight
{
    void initialize(double y_optimum, double width_parameter, ExposureWeight::argument_const_iterator arguments_begin,
                    ExposureWeight::argument_const_iterator arguments_end) override
    {
        // the triangle's half-width matches the Gaussian's full width at half maximum
        ExposureWeight::initialize(y_optimum, width_parameter * FWHM_GAUSSIAN, arguments_begin, arguments_end);
    }

    double weight(double y) override
    {
        const double z = std::fabs(normalize(y));
        return z <= 1.0 ? 1.0 - z : 0.0;
    }
};
