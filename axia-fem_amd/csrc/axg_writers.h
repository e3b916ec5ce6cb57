// Text formatting of results as the reference's DAT and FRD writers print
// them (dat.rs, frd.rs): the number formats and a row formatter that splits
// large tables over host threads and joins the pieces in row order.
#pragma once
#include <cstdint>
#include <functional>
#include <string>

namespace axg {

// frd.rs:7-40: Fortran ES12.5 (sign or space, d.ddddd, E, signed two-digit
// exponent), 12 characters.
void frd_e12(double v, char out[12]);

// Rust's "{:W.PE}" for an f64: P decimals, exponent without '+' or leading
// zeros, right-aligned to width W. Appends to s.
void rust_exp(std::string &s, double v, int prec, int width);

// Rust's "{:W}" for an integer: right-aligned to width W. Appends to s.
void rust_int(std::string &s, int64_t v, int width);

// Rows 0..n-1 formatted by row(i, out) (appending to out), in row order,
// appended to o; the rows are split into contiguous chunks formatted on up to
// `threads` host threads (0: the hardware's count, at most 16).
void format_rows(std::string &o, int64_t n,
                 const std::function<void(int64_t, std::string &)> &row, int threads = 0);

}  // namespace axg
