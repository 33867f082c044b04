// User exposure-weight functions for the GPU path of enfuse.
//
// enfuse weights every pixel by ExposureWeight::weight(y) of its gray value y
// in [0, 1] (src/enfuse.h:623-651); the function is a built-in or a plugin
// object found with dlsym (src/exposure_weight.cc:44-80,
// doc/examples/enfuse/linear.cc).  The GPU never calls host code per pixel:
// for an integer image the weight depends only on the gray code, so the host
// evaluates the function once per code and the kernels look the table up,
// which gives the same weights bit for bit.
//
// This header declares the plugin interface class the reference's
// src/exposure_weight_base.h:29-88 declares (same public members, virtual
// functions in the same order, same data members), so a plugin built against
// either header works with the other.  When the reference's header was
// included first, its declaration is used.
#pragma once

#include <stddef.h>

#include <algorithm>
#include <iterator>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifndef EXPOSURE_WEIGHT_BASE_INCLUDED
#define EXPOSURE_WEIGHT_BASE_INCLUDED

#define EXPOSURE_WEIGHT_INTERFACE_VERSION 2

// Full width at half maximum of the Gaussian exposure weight, 2 sqrt(2 ln 2);
// other weight functions rescale their width to it.
#define FWHM_GAUSSIAN 2.3548200450309493820231386529193992755

class ExposureWeight
{
public:
    typedef std::vector<std::string> argument_list_t;
    typedef argument_list_t::const_iterator argument_const_iterator;

    ExposureWeight() : y_optimum_(0.5), width_(0.2) {}
    ExposureWeight(double y_optimum, double width_parameter) : y_optimum_(y_optimum), width_(width_parameter)
    {
        validate();
    }

    // Sets optimum, width and the plugin's extra arguments; rejects an
    // optimum outside [0, 1] or a width <= 0 with std::invalid_argument.
    virtual void initialize(double y_optimum, double width_parameter, argument_const_iterator arguments_begin,
                            argument_const_iterator arguments_end)
    {
        y_optimum_ = y_optimum;
        width_ = width_parameter;
        arguments_.assign(arguments_begin, arguments_end);
        validate();
    }

    double optimum() const { return y_optimum_; }
    double width() const { return width_; }
    const argument_list_t& arguments() const { return arguments_; }

    // Offset of y from the optimum in units of the width.
    virtual double normalize(double y) const { return (y - optimum()) / width(); }

    // Weight in [0, 1] of a gray value y in [0, 1].
    virtual double weight(double y) = 0;

    virtual ~ExposureWeight() {}

    struct error : public std::runtime_error
    {
        error(const std::string& message) : std::runtime_error(message) {}
    };

    virtual int interface_version() const { return EXPOSURE_WEIGHT_INTERFACE_VERSION; }

private:
    void validate() const
    {
        if (y_optimum_ < 0.0 || y_optimum_ > 1.0) throw std::invalid_argument("y_optimum");
        if (width_ <= 0.0) throw std::invalid_argument("width_parameter");
    }

    double y_optimum_;
    double width_;
    argument_list_t arguments_;
};

#endif  // EXPOSURE_WEIGHT_BASE_INCLUDED

namespace enblend_amd {

// Results of check_weight_function (src/exposure_weight.h:41-42).
enum weight_function_check_t { WEIGHT_OK, WEIGHT_NEGATIVE, WEIGHT_NON_UNIT, WEIGHT_DEGENERATE };

// check_weight_function (src/exposure_weight.cc:237-264): weight(i / (n - 1))
// for i < n must lie in [0, 1], and not all of them may be zero (at most the
// smallest subnormal).  enfuse exits on anything but WEIGHT_OK
// (src/enfuse.cc:1920-1945).
template <class W>
weight_function_check_t check_weight_function(W* fn, int n = 65536)
{
    if (n < 2) throw std::invalid_argument("check_weight_function: n < 2");
    int zeros = 0;
    for (int i = 0; i < n; ++i) {
        const double w = fn->weight(static_cast<double>(i) / static_cast<double>(n - 1));
        if (w < 0.0) return WEIGHT_NEGATIVE;
        if (w > 1.0) return WEIGHT_NON_UNIT;
        if (w >= 0.0 && w <= std::numeric_limits<double>::denorm_min()) ++zeros;
    }
    return zeros == n ? WEIGHT_DEGENERATE : WEIGHT_OK;
}

// weight(code / max) for every gray code of an integer image with `codes`
// levels (256 for 8-bit, 65536 for 16-bit images): the table the GPU path
// looks the exposure weight up in (enbl_enfuse_params.exposure_table), equal
// to what ExposureFunctor computes per pixel (src/enfuse.h:637-640:
// y = gray / NumericTraits<T>::max()).
template <class W>
std::vector<double> exposure_weight_table(W* fn, int codes)
{
    if (codes != 256 && codes != 65536) throw std::invalid_argument("exposure_weight_table: 256 or 65536 codes");
    std::vector<double> t(static_cast<size_t>(codes));
    const double maxv = static_cast<double>(codes - 1);
    for (int g = 0; g < codes; ++g) t[static_cast<size_t>(g)] = fn->weight(static_cast<double>(g) / maxv);
    return t;
}

}  // namespace enblend_amd
