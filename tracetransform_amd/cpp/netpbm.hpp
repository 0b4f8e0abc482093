// netpbm.hpp -- image and table I/O on either side of the transform.
//
// Behaviour follows c++/src/auxiliary.cpp:
//   readnetpbm  ASCII P2 (gray) / P3 (RGB) with '#' comments, one integer
//               matrix per channel (:26-99); a maxval other than 255 and
//               trailing data only warn
//   writepgm    P2, "cols rows", maxval 255, lines broken after 66 chars (:101-125)
//   writecsv    row per line, ", " between columns, default ostream
//               formatting (6 significant digits) (:127-146)
//   gray2mat    value / 255.0 (:148-158), mat2gray value * 255 / max (:160-178)
// Self-contained (no HIP): the readers and writers are usable without a GPU.
#pragma once

#include <climits>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace tracetransform {

// Column-major integer matrix (one netpbm channel).
struct IntMatrix {
    std::ptrdiff_t rows = 0, cols = 0;
    std::vector<int> v;
    IntMatrix() = default;
    IntMatrix(std::ptrdiff_t r, std::ptrdiff_t c) : rows(r), cols(c), v((size_t)r * c, 0) {}
    int &operator()(std::ptrdiff_t r, std::ptrdiff_t c) { return v[(size_t)(r + c * rows)]; }
    int operator()(std::ptrdiff_t r, std::ptrdiff_t c) const { return v[(size_t)(r + c * rows)]; }
};

namespace detail {

inline void skip_blanks_and_comments(std::istream &s)
{
    for (;;) {
        int c = s.peek();
        if (c == '#') {
            s.ignore(INT_MAX, '\n');
        } else if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
            s.get();
        } else {
            return;
        }
    }
}

template <class T> T read_value(std::istream &s)
{
    skip_blanks_and_comments(s);
    T value;
    s >> value;
    if (s.fail())
        throw std::runtime_error("Error processing file");
    return value;
}

}  // namespace detail

inline std::vector<IntMatrix> readnetpbm(const std::string &filename,
                                         std::ostream *warnings = &std::cerr)
{
    std::ifstream in(filename);
    if (!in.good())
        throw std::runtime_error("could not open input file");
    std::stringstream ss;
    ss << in.rdbuf();

    const std::string magic = detail::read_value<std::string>(ss);
    size_t channels;
    if (magic == "P2")
        channels = 1;
    else if (magic == "P3")
        channels = 3;
    else
        throw std::runtime_error("Invalid Netpbm magic");

    const size_t cols = detail::read_value<size_t>(ss);
    const size_t rows = detail::read_value<size_t>(ss);
    std::vector<IntMatrix> data(channels, IntMatrix((std::ptrdiff_t)rows, (std::ptrdiff_t)cols));
    const size_t maxval = detail::read_value<size_t>(ss);
    if (maxval != 255 && warnings)
        *warnings << "Pixels not properly clipped to [0,255]" << std::endl;

    for (size_t r = 0; r < rows; r++)
        for (size_t c = 0; c < cols; c++)
            for (size_t i = 0; i < channels; i++)
                data[i]((std::ptrdiff_t)r, (std::ptrdiff_t)c) =
                    (int)detail::read_value<unsigned int>(ss);

    for (int c = ss.peek(); c == ' ' || c == '\t' || c == '\n' || c == '\r'; c = ss.peek())
        ss.get();
    if (!ss.eof() && warnings)
        *warnings << "Trailing data at end of image file" << std::endl;
    return data;
}

inline void writepgm(const std::string &filename, const IntMatrix &data)
{
    std::ofstream out(filename);
    out << "P2\n" << data.cols << " " << data.rows << "\n" << 255 << "\n";
    std::streamoff line_start = out.tellp();
    for (std::ptrdiff_t r = 0; r < data.rows; r++)
        for (std::ptrdiff_t c = 0; c < data.cols; c++) {
            out << data(r, c);
            if (out.tellp() - line_start > 66) {
                out << "\n";
                line_start = out.tellp();
            } else {
                out << " ";
            }
        }
}

// `at(r, c)` returns the float at row r, column c.
template <class Getter>
inline void writecsv(const std::string &filename, std::ptrdiff_t rows, std::ptrdiff_t cols,
                     Getter at)
{
    std::ofstream out(filename);
    for (std::ptrdiff_t r = 0; r < rows; r++) {
        for (std::ptrdiff_t c = 0; c < cols; c++) {
            out << (float)at(r, c);
            if (c < cols - 1)
                out << ", ";
        }
        out << "\n";
    }
    out << std::flush;
}

// Column-major float buffer from a gray channel, value / 255 in double.
inline std::vector<float> gray2mat(const IntMatrix &in)
{
    std::vector<float> out(in.v.size());
    for (size_t i = 0; i < in.v.size(); i++)
        out[i] = (float)(in.v[i] / 255.0);
    return out;
}

// Gray channel scaled so the maximum maps to 255 (truncating).
inline IntMatrix mat2gray(const float *colmajor, std::ptrdiff_t rows, std::ptrdiff_t cols)
{
    float maximum = 0;
    for (std::ptrdiff_t i = 0; i < rows * cols; i++)
        if (colmajor[i] > maximum)
            maximum = colmajor[i];
    IntMatrix out(rows, cols);
    for (std::ptrdiff_t i = 0; i < rows * cols; i++)
        out.v[(size_t)i] = (int)(colmajor[i] * 255.0 / maximum);
    return out;
}

}  // namespace tracetransform
