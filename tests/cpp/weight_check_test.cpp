// Host-side checks of include/enblend_amd/exposure_weight.hpp (no GPU):
// check_weight_function's four outcomes on functions built to hit each, the
// gray-code table of a plugin-shaped function, and the interface's
// rejection of a bad optimum or width.  Prints one line per check; exit 0
// when all hold.
#include <stdio.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/enblend_amd/exposure_weight.hpp"
#include "linear_weight.hpp"

namespace {

struct Fn : public ExposureWeight {
    double (*f)(double);
    explicit Fn(double (*g)(double)) : f(g) {}
    double weight(double y) override { return f(y); }
};

int failures = 0;

void expect(bool ok, const char* what)
{
    printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    if (!ok) ++failures;
}

}  // namespace

int main()
{
    using namespace enblend_amd;
    Fn neg([](double y) { return y > 0.999 ? -1e-300 : 0.5; });
    Fn big([](double y) { return y < 1e-4 ? std::nextafter(1.0, 2.0) : 0.5; });
    Fn zero([](double y) { return y == 0.5 ? 4.9e-324 : 0.0; });
    Fn one_above([](double y) { return y == 1.0 ? 1e-300 : 0.0; });
    Fn nan_pt([](double y) { return y == 0.0 ? std::nan("") : 0.25; });
    expect(check_weight_function(&neg) == WEIGHT_NEGATIVE, "negative weight near y = 1 -> NEGATIVE");
    expect(check_weight_function(&big) == WEIGHT_NON_UNIT, "weight above 1 near y = 0 -> NON_UNIT");
    expect(check_weight_function(&zero) == WEIGHT_DEGENERATE, "all weights at most denorm_min -> DEGENERATE");
    expect(check_weight_function(&one_above) == WEIGHT_OK, "one weight above denorm_min -> OK");
    expect(check_weight_function(&nan_pt) == WEIGHT_OK, "NaN is neither negative, above one nor zero -> OK");
    // only points i / (n - 1) are sampled: a negative weight between them passes
    Fn between([](double y) { return (y > 16384.0 / 65535.0 + 1e-9 && y < 16385.0 / 65535.0 - 1e-9) ? -1.0 : 0.5; });
    expect(check_weight_function(&between, 65536) == WEIGHT_OK, "weights between the sample points are not seen");
    LinearWeight lin;
    const std::vector<std::string> args;
    lin.initialize(0.5, 0.2, args.begin(), args.end());
    expect(check_weight_function(&lin) == WEIGHT_OK, "linear plugin -> OK");
    const std::vector<double> t = exposure_weight_table(&lin, 256);
    bool same = t.size() == 256;
    for (int g = 0; same && g < 256; ++g) {
        const double z = std::fabs((g / 255.0 - 0.5) / (0.2 * FWHM_GAUSSIAN));
        same = t[g] == (z <= 1.0 ? 1.0 - z : 0.0);
    }
    expect(same, "8-bit table = weight(g / 255) for every code");
    expect(exposure_weight_table(&lin, 65536)[65535] == lin.weight(1.0), "16-bit table ends at weight(1)");
    bool threw = false;
    try {
        lin.initialize(1.5, 0.2, args.begin(), args.end());
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    expect(threw, "optimum outside [0, 1] rejected");
    threw = false;
    try {
        lin.initialize(0.5, 0.0, args.begin(), args.end());
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    expect(threw, "width 0 rejected");
    expect(lin.interface_version() == EXPOSURE_WEIGHT_INTERFACE_VERSION && EXPOSURE_WEIGHT_INTERFACE_VERSION == 2,
           "interface version 2");
    return failures ? 1 : 0;
}
