// tracetransform.hpp -- C++ drop-in for the reference's sinogram/circus API.
//
// Mirrors the reference's calling surface over the C ABI of include/tt.h:
//   enums and wrappers        c++/src/sinogram.hpp:22-48, c++/src/circus.hpp:27-57
//   operator>> grammars       c++/src/sinogram.cpp:34-65, c++/src/circus.cpp:33-66
//   getSinograms              c++/src/sinogram.hpp:57-59
//   getCircusFunctions        c++/src/circus.hpp:69-71
//   nearest_orthonormal_sinogram  c++/src/circus.hpp:66-67
//   pad, resize, zscore       c++/src/auxiliary.hpp:46-58
//   Transformer               c++/src/transform.hpp:24-38, c++/src/transform.cpp:28-112
// Matrices are column-major like Eigen::MatrixXf (element (r, c) at
// r + c * rows), so reference buffers map one to one.  Errors: invalid
// functional strings throw std::invalid_argument (the reference throws
// boost::program_options::validation_error); failures of the device path throw
// tracetransform::Error with the ABI's message.
#pragma once

#include "netpbm.hpp"
#include "tt.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <istream>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace tracetransform {

struct Error : std::runtime_error {
    explicit Error(const std::string &what) : std::runtime_error(what) {}
};

inline void check(tt_status s, const char *fn)
{
    if (s != TT_OK)
        throw Error(std::string(fn) + ": " + tt_last_error());
}

inline void check_hip(hipError_t e, const char *fn)
{
    if (e != hipSuccess)
        throw Error(std::string(fn) + ": " + hipGetErrorString(e));
}

// Column-major float matrix with the parts of Eigen::MatrixXf the path uses.
class Matrix {
  public:
    Matrix() = default;
    Matrix(std::ptrdiff_t rows, std::ptrdiff_t cols)
        : rows_(rows), cols_(cols), v_((size_t)rows * cols, 0.0f) {}
    static Matrix Zero(std::ptrdiff_t rows, std::ptrdiff_t cols) { return Matrix(rows, cols); }
    std::ptrdiff_t rows() const { return rows_; }
    std::ptrdiff_t cols() const { return cols_; }
    std::ptrdiff_t size() const { return rows_ * cols_; }
    float *data() { return v_.data(); }
    const float *data() const { return v_.data(); }
    float &operator()(std::ptrdiff_t r, std::ptrdiff_t c) { return v_[(size_t)(r + c * rows_)]; }
    float operator()(std::ptrdiff_t r, std::ptrdiff_t c) const { return v_[(size_t)(r + c * rows_)]; }

  private:
    std::ptrdiff_t rows_ = 0, cols_ = 0;
    std::vector<float> v_;
};

using Vector = std::vector<float>;

// ---- functionals -----------------------------------------------------------

enum class TFunctional { Radon, T1, T2, T3, T4, T5, T6, T7 };

struct TFunctionalArguments {};

struct TFunctionalWrapper {
    TFunctionalWrapper() : name("invalid"), functional(TFunctional()) {}
    TFunctionalWrapper(const std::string &_name, const TFunctional &_functional,
                       const TFunctionalArguments &_arguments = TFunctionalArguments())
        : name(_name), functional(_functional), arguments(_arguments) {}
    std::string name;
    TFunctional functional;
    TFunctionalArguments arguments;
};

enum class PFunctional { Hermite, P1, P2, P3 };

struct PFunctionalArguments {
    PFunctionalArguments(std::optional<unsigned int> _order = std::nullopt,
                         std::optional<size_t> _center = std::nullopt)
        : order(_order), center(_center) {}
    std::optional<unsigned int> order;
    std::optional<size_t> center;
};

struct PFunctionalWrapper {
    PFunctionalWrapper() : name("invalid"), functional(PFunctional()) {}
    PFunctionalWrapper(const std::string &_name, const PFunctional &_functional,
                       const PFunctionalArguments &_arguments = PFunctionalArguments())
        : name(_name), functional(_functional), arguments(_arguments) {}
    std::string name;
    PFunctional functional;
    PFunctionalArguments arguments;
};

inline std::istream &operator>>(std::istream &in, TFunctionalWrapper &w)
{
    static const char *names[] = {"Radon", "T1", "T2", "T3", "T4", "T5", "T6", "T7"};
    in >> w.name;
    if (w.name.size() == 1 && w.name[0] >= '0' && w.name[0] <= '7') {
        const int k = w.name[0] - '0';
        w.name = names[k];
        w.functional = (TFunctional)k;
        return in;
    }
    throw std::invalid_argument("invalid T-functional '" + w.name + "'");
}

inline std::istream &operator>>(std::istream &in, PFunctionalWrapper &w)
{
    in >> w.name;
    if (w.name == "1" || w.name == "2" || w.name == "3") {
        w.functional = (PFunctional)(w.name[0] - '0');
        w.name = "P" + w.name;
        return in;
    }
    if (!w.name.empty() && w.name[0] == 'H') {
        w.functional = PFunctional::Hermite;
        if (w.name.size() < 2)
            throw std::invalid_argument("missing order parameter for Hermite P-functional");
        const std::string digits = w.name.substr(1);
        if (digits.find_first_not_of("0123456789") != std::string::npos)
            throw std::invalid_argument("unparseable order parameter for Hermite P-functional");
        w.arguments.order = (unsigned int)std::stoul(digits);
        return in;
    }
    throw std::invalid_argument("invalid P-functional '" + w.name + "'");
}

// ---- logging threshold (c++/src/logger.hpp levels) -------------------------

// A message of level l is shown when l <= threshold (default info).
enum class LogLevel { fatal = -3, error, warning, info = 0, debug, trace };

inline LogLevel &log_threshold()
{
    static LogLevel level = LogLevel::info;
    return level;
}

inline bool log_enabled(LogLevel l) { return (int)l <= (int)log_threshold(); }

// ---- context and device memory ----------------------------------------------

// One context per host thread, bound to the HIP default stream so that the
// blocking hipMemcpy calls below are ordered after the library's kernels.
inline tt_context *context()
{
    static thread_local struct Holder {
        tt_context *ctx = nullptr;
        ~Holder()
        {
            if (ctx)
                tt_context_destroy(ctx);
        }
    } h;
    if (!h.ctx) {
        int device = 0;
        check_hip(hipGetDevice(&device), "hipGetDevice");
        check(tt_context_create(device, &h.ctx), "tt_context_create");
        check(tt_context_set_stream(h.ctx, nullptr), "tt_context_set_stream");
    }
    return h.ctx;
}

// Growable device array of T.
template <class T> class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(size_t count) { reserve(count); }
    ~DeviceArray() { (void)hipFree(p_); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), cap_(o.cap_) { o.p_ = nullptr, o.cap_ = 0; }
    DeviceArray &operator=(DeviceArray &&o) noexcept
    {
        if (this != &o) {
            (void)hipFree(p_);
            p_ = o.p_, cap_ = o.cap_;
            o.p_ = nullptr, o.cap_ = 0;
        }
        return *this;
    }
    void reserve(size_t count)
    {
        if (count <= cap_ && p_)
            return;
        (void)hipFree(p_);
        p_ = nullptr;
        cap_ = 0;
        check_hip(hipMalloc(&p_, sizeof(T) * (count ? count : 1)), "hipMalloc");
        cap_ = count;
    }
    T *get() const { return p_; }
    void upload(const T *host, size_t count)
    {
        reserve(count);
        check_hip(hipMemcpy(p_, host, sizeof(T) * count, hipMemcpyHostToDevice), "hipMemcpy");
    }
    void download(T *host, size_t count, size_t offset = 0) const
    {
        check_hip(hipMemcpy(host, p_ + offset, sizeof(T) * count, hipMemcpyDeviceToHost),
                  "hipMemcpy");
    }

  private:
    T *p_ = nullptr;
    size_t cap_ = 0;
};

inline std::vector<int32_t> tcodes(const std::vector<TFunctionalWrapper> &tfunctionals)
{
    std::vector<int32_t> codes;
    for (const auto &t : tfunctionals)
        codes.push_back((int32_t)t.functional);
    return codes;
}

// Hermite entries need order and center (the reference dereferences both).
inline std::vector<tt_pfunc> pfuncs(const std::vector<PFunctionalWrapper> &pfunctionals)
{
    std::vector<tt_pfunc> pf;
    for (const auto &p : pfunctionals) {
        tt_pfunc f = {(int32_t)p.functional, -1, -1};
        if (p.functional == PFunctional::Hermite) {
            if (!p.arguments.order || !p.arguments.center)
                throw std::invalid_argument("Hermite P-functional without order or center");
            f.order = (int32_t)*p.arguments.order;
            f.center = (int32_t)*p.arguments.center;
        }
        pf.push_back(f);
    }
    return pf;
}

inline int angle_count(unsigned int angle_stepsize)
{
    if (angle_stepsize == 0)
        throw std::invalid_argument("angle stepsize must be positive");
    return (int)(360 / angle_stepsize);
}

// ---- the path --------------------------------------------------------------

// c++/src/sinogram.hpp:57-59.  `input` is the padded square image.
inline std::vector<Matrix> getSinograms(const Matrix &input, unsigned int angle_stepsize,
                                        const std::vector<TFunctionalWrapper> &tfunctionals)
{
    if (input.rows() != input.cols())
        throw std::invalid_argument("getSinograms expects a square (padded) image");
    const int n = (int)input.rows(), A = angle_count(angle_stepsize);
    const std::vector<int32_t> codes = tcodes(tfunctionals);
    std::vector<float> out(codes.size() * (size_t)n * A);
    check(tt_sinograms_host(context(), input.data(), n, A, (float)angle_stepsize, codes.data(),
                            (int)codes.size(), out.data()),
          "getSinograms");
    std::vector<Matrix> sinograms;
    for (size_t t = 0; t < codes.size(); t++) {
        Matrix m(n, A);
        std::copy(out.begin() + t * (size_t)n * A, out.begin() + (t + 1) * (size_t)n * A,
                  m.data());
        sinograms.push_back(std::move(m));
    }
    return sinograms;
}

// c++/src/circus.hpp:69-71.
inline std::vector<Vector> getCircusFunctions(const Matrix &input,
                                              const std::vector<PFunctionalWrapper> &pfunctionals)
{
    const std::vector<tt_pfunc> pf = pfuncs(pfunctionals);
    const int n = (int)input.rows(), A = (int)input.cols();
    std::vector<float> out(pf.size() * (size_t)A);
    check(tt_circus_host(context(), input.data(), n, A, pf.data(), (int)pf.size(), out.data()),
          "getCircusFunctions");
    std::vector<Vector> circus;
    for (size_t p = 0; p < pf.size(); p++)
        circus.emplace_back(out.begin() + p * A, out.begin() + (p + 1) * A);
    return circus;
}

// c++/src/circus.hpp:66-67.
inline Matrix nearest_orthonormal_sinogram(const Matrix &input, size_t &new_center)
{
    const int n = (int)input.rows(), A = (int)input.cols();
    const int cap = 3 * n;
    tt_context *ctx = context();
    DeviceArray<float> d_in, d_out((size_t)cap * A);
    d_in.upload(input.data(), (size_t)n * A);
    int rows = 0, center = 0;
    check(tt_nearest_orthonormal_sinogram(ctx, d_in.get(), n, A, d_out.get(), cap, &rows,
                                          &center),
          "nearest_orthonormal_sinogram");
    Matrix nos(rows, A);
    d_out.download(nos.data(), (size_t)rows * A);
    new_center = (size_t)center;
    return nos;
}

// c++/src/auxiliary.hpp:52 (pad), computed on the host.
inline Matrix pad(const Matrix &image)
{
    int n, dfx, dfy;
    check(tt_pad_geometry((int)image.rows(), (int)image.cols(), &n, &dfx, &dfy), "pad");
    Matrix out = Matrix::Zero(n, n);
    for (std::ptrdiff_t c = 0; c < image.cols(); c++)
        for (std::ptrdiff_t r = 0; r < image.rows(); r++)
            out(r + dfy, c + dfx) = image(r, c);
    return out;
}

// c++/src/auxiliary.hpp:46 (resize) on the device.
inline Matrix resize(const Matrix &input, size_t rows, size_t cols)
{
    tt_context *ctx = context();
    DeviceArray<float> d_in, d_out(rows * cols);
    d_in.upload(input.data(), (size_t)input.size());
    check(tt_resize(ctx, d_in.get(), (int)input.rows(), (int)input.cols(), d_out.get(), (int)rows,
                    (int)cols),
          "resize");
    Matrix out((std::ptrdiff_t)rows, (std::ptrdiff_t)cols);
    d_out.download(out.data(), rows * cols);
    return out;
}

// c++/src/auxiliary.cpp:305-318.  A single element gives NaN (0/0 variance).
inline Vector zscore(const Vector &input)
{
    if (input.size() < 2)
        return input.empty() ? Vector() : Vector(input.size(), NAN);
    tt_context *ctx = context();
    DeviceArray<float> d_in, d_out(input.size());
    d_in.upload(input.data(), input.size());
    check(tt_zscore(ctx, d_in.get(), (int)input.size(), 1, d_out.get()), "zscore");
    Vector out(input.size());
    d_out.download(out.data(), input.size());
    return out;
}

// c++/src/transform.hpp:24-38.  The constructor stretches (orthonormal mode)
// and pads on the device and keeps the padded image resident; getTransform
// keeps sinograms and circus functions on the device and copies back only
// the signatures, plus the sinograms when it writes them.
class Transformer {
  public:
    Transformer(const Matrix &image, const std::string &basename, unsigned int angle_stepsize,
                bool orthonormal)
        : _basename(basename), _orthonormal(orthonormal), _angle_stepsize(angle_stepsize)
    {
        tt_context *ctx = context();
        int rows = (int)image.rows(), cols = (int)image.cols();
        DeviceArray<float> d_src;
        d_src.upload(image.data(), (size_t)image.size());
        if (_orthonormal) {
            const int ndiag = (int)std::ceil(360.0 / angle_stepsize);
            const int nsize = (int)std::ceil(ndiag / std::sqrt(2));
            DeviceArray<float> d_resized((size_t)nsize * nsize);
            check(tt_resize(ctx, d_src.get(), rows, cols, d_resized.get(), nsize, nsize),
                  "resize");
            d_src = std::move(d_resized);
            rows = cols = nsize;
        }
        int dfx, dfy;
        check(tt_pad_geometry(rows, cols, &_n, &dfx, &dfy), "pad");
        _d_image.reserve((size_t)_n * _n);
        check(tt_pad(ctx, d_src.get(), rows, cols, _d_image.get()), "pad");
        check(tt_context_synchronize(ctx), "tt_context_synchronize");
    }

    // Padded image, n x n.
    Matrix image() const
    {
        Matrix m(_n, _n);
        _d_image.download(m.data(), (size_t)_n * _n);
        return m;
    }

    // Signatures, A x (nt * np): column t * np + p is the z-score of circus
    // function p of sinogram t.  Orthonormal mode writes each Hermite
    // centre back into `pfunctionals` (c++/src/transform.cpp:78-84).  With
    // `write_data`, <basename>.csv receives the signatures and, at debug
    // logging, <basename>-<T>.csv/.pgm each sinogram.
    Matrix getTransform(const std::vector<TFunctionalWrapper> &tfunctionals,
                        std::vector<PFunctionalWrapper> &pfunctionals,
                        bool write_data = false) const
    {
        tt_context *ctx = context();
        const int n = _n, A = angle_count(_angle_stepsize);
        const int nt = (int)tfunctionals.size(), np = (int)pfunctionals.size();
        Matrix signatures(A, (std::ptrdiff_t)nt * np);
        if (nt == 0)
            return signatures;
        const std::vector<int32_t> codes = tcodes(tfunctionals);
        _d_sino.reserve((size_t)nt * A * n);
        check(tt_sinograms(ctx, _d_image.get(), n, A, (float)_angle_stepsize, codes.data(), nt,
                           _d_sino.get(), nullptr),
              "getSinograms");
        if (write_data && log_enabled(LogLevel::debug))
            write_sinograms(tfunctionals, n, A);
        if (np == 0)
            return signatures;
        _d_circus.reserve((size_t)nt * np * A);
        if (!_orthonormal) {
            const std::vector<tt_pfunc> pf = pfuncs(pfunctionals);
            check(tt_circus(ctx, _d_sino.get(), n, A, nt, pf.data(), np, _d_circus.get(),
                            nullptr),
                  "getCircusFunctions");
        } else {
            // every sinogram's nearest orthonormal sinogram in one batched solve
            const int cap = 3 * n;
            _d_nos.reserve((size_t)nt * cap * A);
            std::vector<int> rows(nt), centers(nt);
            check(tt_nearest_orthonormal_sinograms(ctx, _d_sino.get(), n, A, nt, _d_nos.get(),
                                                   cap, rows.data(), centers.data()),
                  "nearest_orthonormal_sinogram");
            for (int t = 0; t < nt; t++) {
                for (auto &p : pfunctionals)
                    if (p.functional == PFunctional::Hermite)
                        p.arguments.center = (size_t)centers[t];
                const std::vector<tt_pfunc> pf = pfuncs(pfunctionals);
                check(tt_circus(ctx, _d_nos.get() + (size_t)t * cap * A, rows[t], A, 1,
                                pf.data(), np, _d_circus.get() + (size_t)t * np * A, nullptr),
                      "getCircusFunctions");
            }
        }
        if (A >= 2) {
            _d_sig.reserve((size_t)nt * np * A);
            check(tt_zscore(ctx, _d_circus.get(), A, nt * np, _d_sig.get()), "zscore");
            _d_sig.download(signatures.data(), (size_t)nt * np * A);
        } else {
            for (std::ptrdiff_t i = 0; i < signatures.size(); i++)
                signatures.data()[i] = NAN;
        }
        if (write_data)
            writecsv(_basename + ".csv", signatures.rows(), signatures.cols(),
                     [&](std::ptrdiff_t r, std::ptrdiff_t c) { return signatures(r, c); });
        return signatures;
    }

  private:
    void write_sinograms(const std::vector<TFunctionalWrapper> &tfunctionals, int n, int A) const
    {
        std::vector<float> host((size_t)n * A);
        for (size_t t = 0; t < tfunctionals.size(); t++) {
            _d_sino.download(host.data(), host.size(), t * host.size());
            const std::string stem = _basename + "-" + tfunctionals[t].name;
            writecsv(stem + ".csv", n, A,
                     [&](std::ptrdiff_t r, std::ptrdiff_t c) { return host[(size_t)(r + c * n)]; });
            writepgm(stem + ".pgm", mat2gray(host.data(), n, A));
        }
    }

    std::string _basename;
    bool _orthonormal;
    unsigned int _angle_stepsize;
    int _n = 0;
    DeviceArray<float> _d_image;
    mutable DeviceArray<float> _d_sino, _d_circus, _d_nos, _d_sig;
};

}  // namespace tracetransform
