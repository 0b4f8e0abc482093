// tt_demo -- command-line front end with the reference demo's options and
// outputs (c++/src/demo.cpp:52-257), running the transform on the GPU.
//
//   tt_demo -m calculate|profile|benchmark -T <t> [-T <t> ...] [-P <p> ...]
//           [-a <step>] [-n <iterations>] [-q|-v|-d] <image.pgm|image.ppm> ...
//
// calculate writes <base>.csv (signatures) and, with -v or -d, each sinogram
// as <base>-<T>.csv and <base>-<T>.pgm (c++/src/transform.cpp:61-72,107-111);
// profile runs the transform without output; benchmark runs one warm-up and
// then prints t_<i>=<seconds> per timed transform (c++/src/demo.cpp:222-248).
// A P3 (RGB) image is processed per channel as <base>_c1 .. <base>_c3.
// Exit status 1 on invalid usage or input, with the message on stderr.
#include "tracetransform.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace tt = tracetransform;

namespace {

enum class Mode { none, calculate, profile, benchmark };

const char *kUsage =
    "Allowed options:\n"
    "  -h [ --help ]               produce help message\n"
    "  -q [ --quiet ]              only display errors and warnings\n"
    "  -v [ --verbose ]            display some more details\n"
    "  -d [ --debug ]              display even more details\n"
    "  -T [ --t-functional ] arg   T-functionals (0..7)\n"
    "  -P [ --p-functional ] arg   P-functionals (1..3, H<order>)\n"
    "  -a [ --angle ] arg (=1)     angle stepsize\n"
    "  -m [ --mode ] arg           execution mode ('calculate', 'profile' or 'benchmark')\n"
    "  -n [ --iterations ] arg     amount of iterations to run\n"
    "  -i [ --inputs ] arg         images to process\n";

struct UsageError : std::runtime_error {
    explicit UsageError(const std::string &w) : std::runtime_error(w) {}
};

struct Options {
    bool help = false, quiet = false, verbose = false, debug = false;
    std::vector<tt::TFunctionalWrapper> tfunctionals;
    std::vector<tt::PFunctionalWrapper> pfunctionals;
    unsigned int angle = 1;
    Mode mode = Mode::none;
    bool have_iterations = false;
    unsigned int iterations = 0;
    std::vector<std::string> inputs;
};

unsigned int parse_unsigned(const std::string &v, const std::string &opt)
{
    if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
        throw UsageError("the argument ('" + v + "') for option '" + opt + "' is invalid");
    return (unsigned int)std::stoul(v);
}

Options parse(int argc, char **argv)
{
    Options o;
    struct Spec {
        const char *shrt, *lng;
        bool takes_value;
    };
    static const Spec specs[] = {
        {"h", "help", false},         {"q", "quiet", false},       {"v", "verbose", false},
        {"d", "debug", false},        {"T", "t-functional", true}, {"P", "p-functional", true},
        {"a", "angle", true},         {"m", "mode", true},         {"n", "iterations", true},
        {"i", "inputs", true},
    };
    for (int k = 1; k < argc; k++) {
        std::string arg = argv[k];
        const Spec *spec = nullptr;
        std::string value;
        bool inline_value = false;
        if (arg.size() > 2 && arg.compare(0, 2, "--") == 0) {
            std::string name = arg.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                inline_value = true;
            }
            for (const Spec &s : specs)
                if (name == s.lng)
                    spec = &s;
            if (!spec)
                throw UsageError("unrecognised option '" + arg + "'");
        } else if (arg.size() >= 2 && arg[0] == '-' && arg[1] != '-') {
            for (const Spec &s : specs)
                if (arg[1] == s.shrt[0])
                    spec = &s;
            if (!spec)
                throw UsageError("unrecognised option '" + arg + "'");
            if (arg.size() > 2) {
                value = arg.substr(2);
                inline_value = true;
            }
        } else {
            o.inputs.push_back(arg);
            continue;
        }
        if (!spec->takes_value) {
            if (inline_value)
                throw UsageError("option '" + std::string(spec->lng) + "' does not take any arguments");
        } else if (!inline_value) {
            if (k + 1 >= argc)
                throw UsageError("the required argument for option '--" + std::string(spec->lng) +
                                 "' is missing");
            value = argv[++k];
        }
        const std::string lng = spec->lng;
        if (lng == "help") {
            o.help = true;
        } else if (lng == "quiet") {
            o.quiet = true;
        } else if (lng == "verbose") {
            o.verbose = true;
        } else if (lng == "debug") {
            o.debug = true;
        } else if (lng == "t-functional") {
            tt::TFunctionalWrapper w;
            std::istringstream in(value);
            try {
                in >> w;
            } catch (const std::invalid_argument &) {
                throw UsageError("the argument ('" + value + "') for option '--t-functional' is invalid");
            }
            o.tfunctionals.push_back(w);
        } else if (lng == "p-functional") {
            tt::PFunctionalWrapper w;
            std::istringstream in(value);
            try {
                in >> w;
            } catch (const std::invalid_argument &e) {
                throw UsageError("the argument ('" + value + "') for option '--p-functional' is invalid: " +
                                 e.what());
            }
            o.pfunctionals.push_back(w);
        } else if (lng == "angle") {
            o.angle = parse_unsigned(value, "--angle");
        } else if (lng == "mode") {
            if (value == "calculate")
                o.mode = Mode::calculate;
            else if (value == "profile")
                o.mode = Mode::profile;
            else if (value == "benchmark")
                o.mode = Mode::benchmark;
            else
                throw UsageError("the argument ('" + value + "') for option '--mode' is invalid");
        } else if (lng == "iterations") {
            o.iterations = parse_unsigned(value, "--iterations");
            o.have_iterations = true;
        } else if (lng == "inputs") {
            o.inputs.push_back(value);
        }
    }
    return o;
}

std::string stem_of(const std::string &path)
{
    const size_t slash = path.find_last_of('/');
    std::string file = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = file.find_last_of('.');
    if (dot != std::string::npos && dot != 0)
        file = file.substr(0, dot);
    return file;
}

std::string lower_extension(const std::string &path)
{
    const size_t slash = path.find_last_of('/');
    const std::string file = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = file.find_last_of('.');
    std::string ext = (dot == std::string::npos || dot == 0) ? "" : file.substr(dot);
    for (char &c : ext)
        c = (char)std::tolower((unsigned char)c);
    return ext;
}

bool file_exists(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (f)
        std::fclose(f);
    return f != nullptr;
}

int run(const Options &o)
{
    // c++/src/demo.cpp:153-169: all P-functionals Hermite, or none.
    size_t hermite = 0;
    for (const auto &p : o.pfunctionals)
        hermite += p.functional == tt::PFunctional::Hermite;
    if (hermite != 0 && hermite != o.pfunctionals.size()) {
        std::cerr << "error: Cannot mix regular and orthonormal P-functionals" << std::endl;
        return 1;
    }
    const bool orthonormal = hermite != 0;
    if (o.angle == 0) {
        std::cerr << "error: the angle stepsize must be positive" << std::endl;
        return 1;
    }
    if (o.mode == Mode::benchmark && !o.have_iterations) {
        std::cerr << "Invalid usage: the option '--iterations' is required but missing"
                  << std::endl;
        return 1;
    }

    std::vector<tt::PFunctionalWrapper> pfunctionals = o.pfunctionals;
    for (const std::string &input : o.inputs) {
        if (!file_exists(input)) {
            std::cerr << "error: Input file does not exist: " << input << std::endl;
            return 1;
        }
        const std::string ext = lower_extension(input);
        if (ext != ".pgm" && ext != ".ppm") {
            std::cerr << "error: Unrecognized input file format: " << input << std::endl;
            return 1;
        }
        std::vector<tt::IntMatrix> components;
        try {
            components = tt::readnetpbm(input, &std::cerr);
        } catch (const std::runtime_error &e) {
            std::cerr << "error: " << input << ": " << e.what() << std::endl;
            return 1;
        }
        const std::string basename = stem_of(input);
        for (size_t c = 0; c < components.size(); c++) {
            const std::string name =
                components.size() == 1 ? basename : basename + "_c" + std::to_string(c + 1);
            const tt::IntMatrix &gray = components[c];
            tt::Matrix image(gray.rows, gray.cols);
            const std::vector<float> scaled = tt::gray2mat(gray);
            std::copy(scaled.begin(), scaled.end(), image.data());
            tt::Transformer transformer(image, name, o.angle, orthonormal);
            if (o.mode == Mode::calculate) {
                transformer.getTransform(o.tfunctionals, pfunctionals, true);
            } else if (o.mode == Mode::profile) {
                transformer.getTransform(o.tfunctionals, pfunctionals, false);
            } else {
                transformer.getTransform(o.tfunctionals, pfunctionals, false);
                for (unsigned int i = 0; i < o.iterations; i++) {
                    const auto t0 = std::chrono::high_resolution_clock::now();
                    transformer.getTransform(o.tfunctionals, pfunctionals, false);
                    const auto t1 = std::chrono::high_resolution_clock::now();
                    if (tt::log_enabled(tt::LogLevel::info))
                        std::cout << "t_" << i + 1 << "="
                                  << std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0)
                                             .count() /
                                         1000000.0
                                  << std::endl;
                }
            }
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Options o;
    try {
        o = parse(argc, argv);
    } catch (const UsageError &e) {
        std::cerr << "Error " << e.what() << std::endl;
        std::cout << kUsage << std::endl;
        return 1;
    }
    if (o.help) {
        std::cout << kUsage << std::endl;
        return 0;
    }
    std::string missing;
    if (o.tfunctionals.empty())
        missing = "--t-functional";
    else if (o.mode == Mode::none)
        missing = "--mode";
    else if (o.inputs.empty())
        missing = "--inputs";
    if (!missing.empty()) {
        std::cerr << "Invalid usage: the option '" << missing << "' is required but missing"
                  << std::endl;
        std::cout << kUsage << std::endl;
        return 1;
    }
    if (o.debug)
        tt::log_threshold() = tt::LogLevel::trace;
    else if (o.verbose)
        tt::log_threshold() = tt::LogLevel::debug;
    else if (o.quiet)
        tt::log_threshold() = tt::LogLevel::warning;
    try {
        return run(o);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
